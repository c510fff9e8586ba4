"""Many independent observations imaged together: one captured HIP graph replays the invert and
predict of every observation, with tile side and work partition picked by timing.

ARL images a list of observations by calling invert / predict once per observation
(imaging_serial.py:41-216), rebuilding the convolution mapping every call.  Here each
observation keeps its plan (the mapping and the tile-sorted entry lists) for as long as its
uvw do not change, and the whole sequence of kernels of all observations is captured once into
a graph (torch.cuda.CUDAGraph, a hipGraph on ROCm) and replayed with one launch.  The library
calls on a plan allocate nothing, record no memset and never synchronise after their first
call, which is what makes the capture valid; the first, eager call is the warm-up below.
"""

import ctypes
import json
import math
import os
import time

import torch

from . import _lib
from .imaging import Imager, _shift_lmn
from .data_models import device
from .plan import kernel_struct, stream_handle

__all__ = ["Autotuner", "BatchImager"]

TILES = (32, 64)
CHUNKS = (1024, 4096, 16384)


class Autotuner:
    """Chooses (tile, chunk) per problem shape by timing invert + predict on the problem itself.

    The key is (N, support, oversampling, planes, precision, mode, log2 of the row count), so
    observations of similar shape share a measurement.  Choices persist in ``path`` (JSON) when
    given."""

    def __init__(self, path=None, tiles=TILES, chunks=CHUNKS, reps=3):
        self.path = path
        self.tiles = tuple(tiles)
        self.chunks = tuple(chunks)
        self.reps = reps
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = {k: tuple(v) for k, v in json.load(f).items()}

    @staticmethod
    def key(vis, im, cf, mode, vis_slices):
        n = im.data.shape[-1]
        return "%d/%d/%d/%d/%d/%s/%d" % (n, cf.support, cf.oversampling, vis_slices, vis.precision, mode,
                                          int(math.log2(max(vis.nvis, 1))))

    def _time(self, imager, vis, model, reps):
        out_img = torch.empty((imager.plan.nchan, 1, imager.plan.npixel, imager.plan.npixel),
                              dtype=model.dtype, device=model.device)
        out_vis = torch.empty_like(vis.vis)
        imager.invert(vis, out=out_img)
        imager.predict(vis, model, out=out_vis)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            imager.invert(vis, out=out_img)
            imager.predict(vis, model, out=out_vis)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def choose(self, vis, im, gcfcf, mode, vis_slices, model):
        """(tile, chunk, seconds or None); measures the candidates the first time a key is seen."""
        k = self.key(vis, im, gcfcf[1], mode, vis_slices)
        if k in self.table:
            return self.table[k]
        best = None
        n = im.data.shape[-1]
        for tile in self.tiles:
            if tile > n:
                continue
            for chunk in self.chunks:
                imager = Imager(vis, im, gcfcf=gcfcf, mode=mode, vis_slices=vis_slices, tile=tile, chunk=chunk)
                t = self._time(imager, vis, model, self.reps)
                del imager
                if best is None or t < best[2]:
                    best = (tile, chunk, t)
        self.table[k] = best
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
        return best


class BatchImager:
    """Invert and predict for a list of observations, replayed from one captured graph.

    tasks: list of dicts with keys vis (Visibility), im (Image template), model (image data
    [nchan, 1, N, N] on the device, for predict), gcfcf ((gcf, cf)), and optionally mode
    ("wstack" or "2d"), vis_slices, dopsf, normalize.  After ``run()``, ``images[i]`` holds the
    dirty image of task i, ``predicted[i]`` its model visibilities and ``imagers[i].sumwt()``
    its sum of weights.
    """

    def __init__(self, tasks, autotuner=None, graph=True, persistent=False):
        self.tasks = tasks
        self.imagers = []
        self.images = []
        self.predicted = []
        self.choices = []
        for t in tasks:
            mode = t.get("mode", "wstack")
            slices = t.get("vis_slices", 1)
            tile, chunk = 0, 0
            if autotuner is not None:
                tile, chunk, _ = autotuner.choose(t["vis"], t["im"], t["gcfcf"], mode, slices, t["model"])
            self.choices.append((tile, chunk))
            imager = Imager(t["vis"], t["im"], gcfcf=t["gcfcf"], mode=mode, vis_slices=slices, tile=tile,
                            chunk=chunk)
            self.imagers.append(imager)
            n = imager.plan.npixel
            self.images.append(torch.empty((imager.plan.nchan, 1, n, n), dtype=t["model"].dtype,
                                           device=t["model"].device))
            self.predicted.append(torch.empty_like(t["vis"].vis))
        self.batch = None
        self._eager()  # warm-up: sizes every buffer and caches every kernel table
        torch.cuda.synchronize()
        # persistent: the gridding (and degridding) of every observation of one kernel support in
        # one launch (arlmi_batch_grid / arlmi_batch_degrid); the transforms stay per observation
        if persistent:
            self._make_batch()
        self.graph = None
        if graph:
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._eager()
            torch.cuda.synchronize()

    def _make_batch(self):
        for t in self.tasks:
            if t.get("dopsf", False) != self.tasks[0].get("dopsf", False):
                raise ValueError("a persistent batch grids every observation with the same dopsf")
        items = (_lib.BatchItem * len(self.tasks))()
        self._keep = []
        precision = None
        for i, (t, imager, pv) in enumerate(zip(self.tasks, self.imagers, self.predicted)):
            if _shift_lmn(t["vis"], imager.im) is not None:
                raise ValueError("a persistent batch does not shift visibilities to an image centre")
            precision = imager.precision
            k, keep = kernel_struct(imager.cf, imager.precision)
            self._keep.append(keep)
            it = items[i]
            it.plan = imager.plan.handle.value
            it.d_vis = t["vis"].vis.data_ptr()
            it.d_weight = t["vis"].data["imaging_weight"].data_ptr()
            it.kernel = k
            it.d_grids = imager.grids.data_ptr()
            it.d_sumwt = imager.sumwt_planes.data_ptr()
            it.d_vis_out = pv.data_ptr()
        handle = ctypes.c_void_p()
        _lib.check(_lib.lib().arlmi_batch_create(items, len(self.tasks), precision, ctypes.byref(handle)))
        self.batch = handle

    def _eager(self):
        if self.batch is not None:
            L = _lib.lib()
            dopsf = int(bool(self.tasks[0].get("dopsf", False)))
            _lib.check(L.arlmi_batch_grid(self.batch, dopsf, stream_handle()))
            for t, imager, img in zip(self.tasks, self.imagers, self.images):
                imager.grids_to_image(normalize=t.get("normalize", True), out=img)
            for t, imager in zip(self.tasks, self.imagers):
                imager.image_to_grids(t["model"])
            _lib.check(L.arlmi_batch_degrid(self.batch, stream_handle()))
            return
        for t, imager, img, pv in zip(self.tasks, self.imagers, self.images, self.predicted):
            imager.invert(t["vis"], dopsf=t.get("dopsf", False), normalize=t.get("normalize", True), out=img)
            imager.predict(t["vis"], t["model"], out=pv)

    def __del__(self):
        b = getattr(self, "batch", None)
        if b:
            try:
                _lib.lib().arlmi_batch_destroy(b)
            except Exception:
                pass

    def run(self):
        """Invert and predict every observation (one graph launch, or the eager sequence)."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._eager()

    @property
    def nvis(self):
        return sum(t["vis"].nvis for t in self.tasks)
