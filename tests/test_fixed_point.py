"""The fixed-point rounding of the deterministic gridders (ARLMI_FLAG_DETERMINISTIC), checked on
the host: for a power-of-two scale s and |x s| <= 2^51, the bit pattern of x s + 1.5 * 2^52 minus
that of 1.5 * 2^52 is x s rounded to the nearest integer, ties to even (the doubles in [2^52, 2^53)
are the integers).  The scale rule q = min(60 - e(B), 50 - e(X)), y < 2^e(y), keeps every bounded
sum B 2^q below 2^60 (so 64-bit sums of values within the bound cannot overflow) and every single
tap X 2^q within the rounding's range, while B 2^q stays at or above 2^58 when B is the binding
bound (at most two bits of the integer range unused)."""

import math

import numpy

MAGIC = 1.5 * 2.0 ** 52
MAGIC_BITS = 0x4338000000000000


def _round_bits(x, q):
    y = numpy.asarray(x, dtype=numpy.float64) * numpy.ldexp(1.0, q) + MAGIC
    return y.view(numpy.int64) - MAGIC_BITS


def test_magic_constant():
    assert numpy.array([MAGIC]).view(numpy.int64)[0] == MAGIC_BITS


def test_rounding_identity():
    rng = numpy.random.default_rng(5)
    for q in (-40, -3, 0, 7, 30, 52):
        x = rng.uniform(-1, 1, 100000) * numpy.ldexp(1.0, 51 - q)
        assert numpy.array_equal(_round_bits(x, q), numpy.rint(x * numpy.ldexp(1.0, q)).astype(numpy.int64))
    # ties round to even, the extremes and zero
    x = numpy.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.0, -0.0, 2.0 ** 51, -(2.0 ** 51), 2.0 ** 51 - 0.5])
    assert list(_round_bits(x, 0)) == [0, 2, 2, 0, -2, -2, 0, 0, 2 ** 51, -(2 ** 51), 2 ** 51]


def _exp(y):
    return -1000 if y <= 0 else math.frexp(y)[1]


def _q(b, x):
    return max(-1000, min(1000, min(60 - _exp(b), 50 - _exp(x))))


def test_scale_rule():
    rng = numpy.random.default_rng(6)
    for _ in range(2000):
        x = math.ldexp(rng.uniform(0.5, 1.0), int(rng.integers(-60, 60)))
        n = int(rng.integers(1, 10 ** 9))
        b = x * n * rng.uniform(0.01, 1.0) + x  # a sum bound of at least one tap
        q = _q(b, x)
        assert math.ldexp(b, q) < 2.0 ** 60
        assert math.ldexp(x, q) < 2.0 ** 50
        if 60 - _exp(b) <= 50 - _exp(x):
            assert math.ldexp(b, q) >= 2.0 ** 58
