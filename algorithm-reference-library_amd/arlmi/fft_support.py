"""ARL's shifted FFTs on device arrays (processing_library/fourier_transforms/fft_support.py:21-89)."""

import torch

from . import _lib
from .plan import ptr, stream_handle

__all__ = ["fft", "ifft"]


def _run(a, direction):
    if not torch.is_complex(a):
        a = a.to(torch.complex128 if a.dtype == torch.float64 else torch.complex64)
    n = a.shape[-1]
    if a.shape[-2] != n:
        raise ValueError("square planes only")
    out = a.contiguous().clone()
    precision = 32 if out.dtype == torch.complex64 else 64
    batch = out.numel() // (n * n)
    _lib.check(_lib.lib().arlmi_fft2d(ptr(out), n, batch, direction, precision, stream_handle()))
    return out


def fft(a):
    """Image -> grid: fftshift(fft2(ifftshift(a))) over the last two axes (fft_support.py:21-38)."""
    return _run(a, +1)


def ifft(a):
    """Grid -> image: fftshift(ifft2(ifftshift(a))), 1/N^2 normalised (fft_support.py:55-71)."""
    return _run(a, -1)
