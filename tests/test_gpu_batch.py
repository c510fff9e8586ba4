"""Batched observations replayed from one captured graph, with and without the autotuner, give
the results of eager per-observation invert / predict on every replay."""

import numpy
import pytest

from helpers import rel_err, image_errors

pytestmark = pytest.mark.gpu


def _tasks(nobs=3, seed=31, max_rows=150_000):
    import torch
    import arlmi
    from arlmi.simulation import config3
    obs = config3(nobs=nobs, seed=seed, max_rows=max_rows, min_rows=10_000)
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


@pytest.mark.parametrize("graph", [False, True])
def test_persistent_batch_matches_eager(gpu, graph):
    """One gridding and one degridding launch per kernel support for all observations
    (arlmi_batch_grid / arlmi_batch_degrid), eager and replayed from a graph three times."""
    import torch
    import arlmi
    tasks = _tasks(nobs=4, seed=34)
    ref = _eager_reference(tasks)
    batch = arlmi.BatchImager(tasks, graph=graph, persistent=True)
    assert batch.batch
    for k in range(3):
        for img in batch.images:
            img.fill_(float("nan"))
        for pv in batch.predicted:
            pv.fill_(float("nan"))
        batch.run()
        torch.cuda.synchronize()
        _check(batch, ref)
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


def test_config3_full_batch(gpu):
    """The whole batching workload (64 observations of 10k-4M rows, N from 256 to 2048, S in
    {6, 8, 10}, 8-64 slabs, 102.8 M rows), gridded and degridded with one launch per support and
    replayed from a graph: every observation against its eager run, and the two cheapest
    observations against the oracle."""
    import torch
    import arlmi
    from oracle import arl_oracle as O
    tasks = _tasks(nobs=64, seed=3, max_rows=4_000_000)
    assert sum(t["vis"].nvis for t in tasks) > 90_000_000
    batch = arlmi.BatchImager(tasks, graph=True, persistent=True)
    batch.run()
    torch.cuda.synchronize()
    for i, t in enumerate(tasks):
        imager = arlmi.Imager(t["vis"], t["im"], gcfcf=t["gcfcf"], mode="wstack", vis_slices=t["vis_slices"])
        re, _ = imager.invert(t["vis"])
        pv = imager.predict(t["vis"], t["model"])
        l2, inner, full = image_errors(batch.images[i].cpu().numpy(), re.cpu().numpy())
        assert inner < 1e-5 and full < 5e-3, (i, l2, inner, full)
        assert rel_err(batch.predicted[i].cpu().numpy(), pv.cpu().numpy()) < 1e-5, i
        del imager, re, pv
    cost = [t["im"].data.shape[-1] ** 2 * t["vis_slices"] for t in tasks]
    for i in numpy.argsort(cost)[:2]:
        t = tasks[i]
        n = t["im"].data.shape[-1]
        w = t["im"].wcs.wcs
        d = {"uvw": t["vis"].data["uvw"].cpu().numpy(),
             "vis": t["vis"].vis.cpu().numpy().astype(complex),
             "imaging_weight": t["vis"].data["imaging_weight"].cpu().numpy().astype(float),
             "frequency": t["vis"].data["frequency"].cpu().numpy()}
        g = O.ImageGeom(n, float(t["im"].cellsize), frequency=float(w.crval[3]), channel_bandwidth=float(w.cdelt[3]))
        ogcf, ocf, cfg = O.create_pswf_convolutionfunction(g, 8, t["gcfcf"][1].support)
        rimg, _ = O.invert_list_wstack(d, g, ogcf, ocf, cfg, t["vis_slices"])
        l2, inner, full = image_errors(batch.images[i].cpu().numpy(), rimg)
        assert inner < 1e-5 and full < 5e-3, (i, l2, inner, full)
        model = t["model"].double().cpu().numpy()
        rpv = O.predict_list_wstack(d, model, g, ogcf, ocf, cfg, t["vis_slices"])
        assert rel_err(batch.predicted[i].cpu().numpy(), rpv) < 1e-5, i
