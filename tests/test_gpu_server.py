"""GPU tests of the render server: continuous batching must not change any
result (each request alone == the same request in a mixed stream, bit for
bit), admission by HBM footprint, and the replay accounting."""
import numpy as np
import pytest

import vfray_amd as V
from vfray_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def two_meshes(ctx):
    ms, tfs = [], []
    for k, n in enumerate((14, 20)):
        x, t = V.synth_grid(n, 0x5EED0400 + k, 0)
        ms.append(ctx.mesh(x, t))
        tfs.append([ctx.tf(V.synth_tf(0x5EED0400 + 16 * k + j)) for j in range(2)])
    yield ms, tfs
    for m in ms:
        m.close()


def _requests(tfs, n, sides, seed=0x5EED0410):
    r = WL.Rng(seed, 9)
    reqs = []
    for i in range(n):
        m = i % 2
        d = r.direction()
        side = sides[i % len(sides)]
        v = V.perspective_view(side, side, tfs[m][i % 2], eye=2.5 * d)
        reqs.append((m, v))
    return reqs


def test_server_alone_equals_mix(ctx, two_meshes):
    ms, tfs = two_meshes
    reqs = _requests(tfs, 12, [64, 96, 128, 160])
    print("server mix of 12 requests on 2 meshes", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=2, max_batch=8)
    try:
        tickets = [sv.submit(m, v, keep_image=True) for m, v in reqs]
        got = [sv.wait(t, image=True) for t in tickets]
    finally:
        sv.close()
    assert max(g["batch_size"] for g in got) > 1
    for (m, v), g in zip(reqs, got):
        alone, _ = ctx.render(ms[m], [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        assert g["status"] == 2 and g["latency_ms"] > 0
        assert np.array_equal(g["rgba"], alone[0]["rgba"])


def test_server_rejects_over_budget(ctx, two_meshes):
    ms, tfs = two_meshes
    sv = V.Server(ctx, ms, hbm_budget=64 << 20, streams=2, max_batch=4)
    try:
        v = V.perspective_view(2048, 2048, tfs[0][0], eye=(0, 0, 2.5))
        with pytest.raises(V.VfrError) as e:
            sv.submit(0, v)
        assert e.value.code == -6
        t = sv.submit(0, V.perspective_view(64, 64, tfs[0][0], eye=(0, 0, 2.5)))
        assert sv.wait(t)["status"] == 2
    finally:
        sv.close()


def test_server_replay_report(ctx, two_meshes):
    ms, tfs = two_meshes
    reqs = [V.Request(m, 0, v) for m, v in _requests(tfs, 40, [32, 64, 96])]
    print("server replay of 40 requests", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=2, max_batch=16)
    try:
        burst = sv.replay(reqs, np.zeros(len(reqs)))
        spaced = sv.replay(reqs, np.arange(len(reqs)) * 2e-3)
    finally:
        sv.close()
    rays = sum(r.view.width * r.view.height for r in reqs)
    for rep in (burst, spaced):
        assert rep["requests"] == 40 and rep["rays"] == rays
        assert 0 < rep["latency_p50_ms"] <= rep["latency_p99_ms"] <= rep["latency_max_ms"]
        assert rep["steps"] > 0 and rep["mrays_per_s"] > 0
    assert burst["mean_batch"] > 1.5
    assert burst["steps"] == spaced["steps"]


@pytest.mark.parametrize("composite", [V.COMPOSITE_LINEAR, V.COMPOSITE_LINEAR_OPACITY])
def test_server_linear_compositing(ctx, two_meshes, composite):
    """A server built with a linear compositing mode renders every request
    as the per-view FTB walk does in that mode, bit for bit."""
    ms, tfs = two_meshes
    reqs = _requests(tfs, 6, [64, 96, 128], seed=0x5EED0420 + composite)
    print(f"server linear compositing {composite}: 6 requests", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=2, max_batch=4, composite=composite)
    try:
        tickets = [sv.submit(m, v, keep_image=True) for m, v in reqs]
        got = [sv.wait(t, image=True) for t in tickets]
    finally:
        sv.close()
    for (m, v), g in zip(reqs, got):
        alone, _ = ctx.render(ms[m], [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW, composite=composite)
        ppi, _ = ctx.render(ms[m], [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        assert g["status"] == 2
        assert np.array_equal(g["rgba"], alone[0]["rgba"])
        assert not np.array_equal(g["rgba"], ppi[0]["rgba"])


def test_server_batch_pixel_cap(ctx, two_meshes):
    """max_batch_pixels bounds every batch (a request above the cap still
    runs, alone) and does not change any image."""
    ms, tfs = two_meshes
    reqs = _requests(tfs, 10, [64, 64, 96, 160], seed=0x5EED0430)
    print("server batch pixel cap: 10 requests", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=1, max_batch=16, max_batch_pixels=2 * 64 * 64)
    try:
        tickets = [sv.submit(m, v, keep_image=True) for m, v in reqs]
        got = [sv.wait(t, image=True) for t in tickets]
    finally:
        sv.close()
    for (m, v), g in zip(reqs, got):
        px = v.width * v.height
        assert g["status"] == 2
        assert g["batch_size"] <= (2 if px <= 64 * 64 else 1)
        alone, _ = ctx.render(ms[m], [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        assert np.array_equal(g["rgba"], alone[0]["rgba"])
