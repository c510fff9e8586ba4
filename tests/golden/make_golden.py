"""Writes the golden fixtures of this directory.

arl_known_answers.json: the numbers ARL's own tests assert (kernel peaks and their indices,
coordinate bounds, DFT values), with the test file:line each comes from; written by hand from
those tests, not generated.

golden_small.npz: inputs and FP64 oracle outputs of small seeded problems (2-D invert / predict,
w-stack invert / predict, uniform weighting, BlockVisibility conversion).  Regenerate with
    python tests/golden/make_golden.py
from the repository root; tests/test_golden.py checks that the oracle still reproduces them and
the GPU tests compare the HIP path with them.
"""

import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import arl_oracle as O  # noqa: E402
from helpers import random_rows  # noqa: E402
from test_oracle_vis import small_block  # noqa: E402


def build():
    out = {}
    rng = numpy.random.default_rng(2024)
    # 2-D: 64^2, PSWF S=8 os=8
    n, cell, S = 64, 0.004, 8
    d = random_rows(rng, 400, n, cell, S)
    g = O.ImageGeom(n, cell)
    gcf, cf, cfg = O.create_pswf_convolutionfunction(g, 8, S)
    img, sw = O.invert_2d(d, g, gcf, cf, cfg)
    model = numpy.zeros((1, 1, n, n))
    model[0, 0, rng.integers(16, 48, 5), rng.integers(16, 48, 5)] = rng.uniform(0.5, 2.0, 5)
    pv = O.predict_2d(d, model, g, gcf, cf, cfg)
    for k in ("uvw", "vis", "imaging_weight", "frequency"):
        out["2d_" + k] = d[k]
    out.update({"2d_image": img, "2d_sumwt": sw, "2d_model": model, "2d_predict": pv,
                "2d_geom": numpy.array([n, cell, S, 8])})
    # w-stack: 64^2, 4 slabs
    d = random_rows(rng, 600, n, cell, S, wmax=40.0)
    P = 4
    img, sw = O.invert_list_wstack(d, g, gcf, cf, cfg, P)
    pv = O.predict_list_wstack(d, model, g, gcf, cf, cfg, P)
    for k in ("uvw", "vis", "imaging_weight", "frequency"):
        out["ws_" + k] = d[k]
    out.update({"ws_image": img, "ws_sumwt": sw, "ws_predict": pv, "ws_slices": numpy.array([P])})
    # uniform weighting
    out["wt_new"] = O.weight_visibility(d, g, cfg)
    # BlockVisibility -> rows
    b, *_ = small_block(seed=77, nants=5, ntimes=2, nchan=2)
    rows = O.convert_blocks(b)
    for k, v in b.items():
        out["blk_in_" + k] = v
    for k, v in rows.items():
        out["blk_out_" + k] = v
    return out


if __name__ == "__main__":
    numpy.savez_compressed(os.path.join(HERE, "golden_small.npz"), **build())
    print("wrote", os.path.join(HERE, "golden_small.npz"))
