"""Batched observations replayed from one captured graph, with and without the autotuner, give
the results of eager per-observation invert / predict on every replay."""

import numpy
import pytest

from helpers import rel_err, image_errors

pytestmark = pytest.mark.gpu


def _tasks(nobs=3, seed=31):
    import torch
    import arlmi
    from arlmi.simulation import config3
    obs = config3(nobs=nobs, seed=seed, max_rows=150_000, min_rows=10_000)
    rng = numpy.random.default_rng(seed)
    tasks = []
    for o in obs:
        n = o["im"].data.shape[-1]
        gcf, cf = arlmi.create_pswf_convolutionfunction(o["im"], oversampling=8, support=o["support"])
        model = numpy.zeros((1, 1, n, n))
        model[0, 0, rng.integers(n // 4, 3 * n // 4, 10), rng.integers(n // 4, 3 * n // 4, 10)] = 1.0
        tasks.append({"vis": o["vis"], "im": o["im"], "gcfcf": (gcf, cf), "mode": "wstack",
                      "vis_slices": o["vis_slices"],
                      "model": torch.as_tensor(model).to(device=o["vis"].vis.device, dtype=torch.float32)})
    return tasks


def _eager_reference(tasks):
    import arlmi
    ref = []
    for t in tasks:
        imager = arlmi.Imager(t["vis"], t["im"], gcfcf=t["gcfcf"], mode="wstack", vis_slices=t["vis_slices"])
        re, _ = imager.invert(t["vis"])
        pv = imager.predict(t["vis"], t["model"])
        ref.append((re.cpu().numpy(), pv.cpu().numpy()))
    return ref


def _check(batch, ref):
    # float32 images of two runs differ by the order of the LDS float adds, which the grid
    # correction amplifies up to ~6e4 times at the image corners: inner field 1e-5, full 5e-3
    for img, pv, (rimg, rpv) in zip(batch.images, batch.predicted, ref):
        l2, inner, full = image_errors(img.cpu().numpy(), rimg)
        assert inner < 1e-5 and full < 5e-3, (l2, inner, full)
        assert rel_err(pv.cpu().numpy(), rpv) < 1e-5


def test_graph_replays_match_eager(gpu):
    import torch
    import arlmi
    tasks = _tasks()
    ref = _eager_reference(tasks)
    batch = arlmi.BatchImager(tasks, graph=True)
    for k in range(3):
        for img in batch.images:
            img.fill_(float("nan"))
        batch.run()
        torch.cuda.synchronize()
        _check(batch, ref)
        # eager work between replays
        _eager_reference(tasks[:1])
        torch.cuda.synchronize()


def test_autotuned_batch_matches_eager(gpu, tmp_path):
    import torch
    import arlmi
    tasks = _tasks(nobs=2, seed=32)
    ref = _eager_reference(tasks)
    tuner = arlmi.Autotuner(path=str(tmp_path / "tune.json"), reps=1)
    batch = arlmi.BatchImager(tasks, autotuner=tuner, graph=True)
    for tile, chunk in batch.choices:
        assert tile in arlmi.batch.TILES and chunk in arlmi.batch.CHUNKS
    batch.run()
    torch.cuda.synchronize()
    _check(batch, ref)
    # the table persists and is reused
    again = arlmi.Autotuner(path=str(tmp_path / "tune.json"))
    assert again.table == tuner.table
