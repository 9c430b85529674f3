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


@pytest.mark.parametrize("schedule", [V.SERVE_SLOTS, V.SERVE_PIPELINE])
def test_server_alone_equals_mix(ctx, two_meshes, schedule):
    ms, tfs = two_meshes
    reqs = _requests(tfs, 12, [64, 96, 128, 160])
    print("server mix of 12 requests on 2 meshes", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=2, max_batch=8, schedule=schedule)
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


@pytest.mark.parametrize("schedule", [V.SERVE_SLOTS, V.SERVE_PIPELINE])
def test_server_thin_views_equal_alone(ctx, two_meshes, schedule):
    """Views of partly filled 8x8 tiles (2 pixels wide or tall, odd sides) in
    one batch: the batch's ray list is built per tile, so every request must
    still render as it does alone."""
    ms, tfs = two_meshes
    sizes = [(2, 64), (64, 2), (9, 9), (7, 300), (2, 2), (130, 3), (17, 33), (3, 129)]
    r = WL.Rng(0x5EED0420, 9)
    reqs = []
    for i, (w, h) in enumerate(sizes * 2):
        m = i % 2
        reqs.append((m, V.perspective_view(w, h, tfs[m][i % 2], eye=2.5 * r.direction())))
    print(f"server batch of {len(reqs)} thin views", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=256 << 20, streams=1, max_batch=len(reqs), schedule=schedule)
    try:
        tickets = [sv.submit(m, v, keep_image=True) for m, v in reqs]
        got = [sv.wait(t, image=True) for t in tickets]
    finally:
        sv.close()
    assert max(g["batch_size"] for g in got) > 1
    for (m, v), g in zip(reqs, got):
        alone, _ = ctx.render(ms[m], [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        assert g["status"] == 2
        assert g["width"] == v.width and g["height"] == v.height
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


@pytest.mark.parametrize("schedule", [V.SERVE_SLOTS, V.SERVE_PIPELINE])
def test_server_replay_report(ctx, two_meshes, schedule):
    ms, tfs = two_meshes
    reqs = [V.Request(m, 0, v) for m, v in _requests(tfs, 40, [32, 64, 96])]
    print("server replay of 40 requests", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=2, max_batch=16, schedule=schedule)
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


@pytest.mark.parametrize("schedule", [V.SERVE_SLOTS, V.SERVE_PIPELINE])
def test_server_batch_pixel_cap(ctx, two_meshes, schedule):
    """max_batch_pixels bounds every batch (a request above the cap still
    runs, alone) and does not change any image."""
    ms, tfs = two_meshes
    reqs = _requests(tfs, 10, [64, 64, 96, 160], seed=0x5EED0430)
    print("server batch pixel cap: 10 requests", flush=True)
    sv = V.Server(ctx, ms, hbm_budget=1 << 30, streams=1, max_batch=16, max_batch_pixels=2 * 64 * 64,
                  schedule=schedule)
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


@pytest.mark.parametrize("schedule", [V.SERVE_SLOTS, V.SERVE_PIPELINE])
def test_server_concave_and_dropped_entries(ctx, schedule):
    """Requests on the concave benchmark mesh (96^3 grid, graded, 12 voids:
    pixels with up to 13 entry faces, more than 8 slots hold) render exactly
    as the per-view walk; a request whose rays run along the boundary edges
    of a grid (129 entry faces at a pixel, more than 32) fails with VFR_EMESH
    instead of returning an image with entries dropped, alone and in a
    replay, and the server goes on serving."""
    x, t, tfa, seed = WL.concave()
    mc = ctx.mesh(x, t)
    xg, tg = V.synth_grid(8, 0x5EED0106, 0)
    mg = ctx.mesh(xg, tg)
    tf = ctx.tf(tfa)
    views = WL.sphere_views(seed, 32, [768], tf)[:2]
    print("server: concave 96^3 2x768^2 + an overflowing 2x2 view", flush=True)
    with pytest.raises(V.VfrError, match="dropped"):
        ctx.render(mc, views[:1], mode=V.MODE_FTB, max_entries=8)
    edge = V.ortho_view(2, 2, tf)
    sv = V.Server(ctx, [mc, mg], hbm_budget=2 << 30, streams=2, max_batch=8, schedule=schedule)
    try:
        tickets = [sv.submit(0, v, keep_image=True) for v in views]
        tbad = sv.submit(1, edge, keep_image=True)
        got = [sv.wait(tk, image=True) for tk in tickets]
        with pytest.raises(V.VfrError) as e:
            sv.wait(tbad, image=True)
        assert e.value.code == -8 and "entry" in str(e.value)
        for v, g in zip(views, got):
            alone, _ = ctx.render(mc, [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
            assert g["status"] == 2
            for key in ("rgba",):
                assert np.array_equal(g[key], alone[0][key])
        # a replay that contains the failing request reports it after reaping all
        reqs = [V.Request(0, 0, views[0]), V.Request(1, 0, edge), V.Request(0, 0, views[1])]
        with pytest.raises(V.VfrError) as e:
            sv.replay(reqs, np.zeros(3))
        assert e.value.code == -8
        t2 = sv.submit(0, views[1], keep_image=True)
        g2 = sv.wait(t2, image=True)
        assert np.array_equal(g2["rgba"], got[1]["rgba"])
    finally:
        sv.close()
        mc.close()
        mg.close()


@pytest.mark.parametrize("schedule", [V.SERVE_SLOTS, V.SERVE_PIPELINE])
def test_server_c5_stream_matches_per_view_and_oracle(ctx, schedule):
    """The serving workload itself: its 8 resident meshes (32^3..160^3
    grids, 0.2M-24.6M tets, 4 TFs each) and the first 200 requests of its
    stream, all queued at once.  Every image equals the same request rendered
    alone by the per-view walk, bit for bit, and meets the oracle bar (entry
    faces and step counts bit-exact with the CPU oracle, RGBA within 2e-3);
    for one request of at most 512^2 per mesh also PSNR >= 50 dB against the
    oracle's FP64 re-integration."""
    import oracle as O
    from helpers import oracle_ftb, psnr
    meshes, tf_ids, host = [], [], []
    for (x, t, tfs) in WL.serve_meshes():
        meshes.append(ctx.mesh(x, t))
        tf_ids.append([ctx.tf(tf) for tf in tfs])
        host.append((x, t, tfs))
    print(f"server C5 schedule {schedule}: 8 meshes built", flush=True)
    reqs = WL.serve_requests(200, len(meshes), tf_ids)
    sv = V.Server(ctx, meshes, hbm_budget=16 << 30, streams=2, max_batch=256, max_batch_pixels=4 << 20,
                  schedule=schedule)
    checked = set()
    inputs = {}
    psi = V.psi_generate()
    try:
        tickets = [sv.submit(r.mesh, r.view, keep_image=True) for r in reqs]
        sizes = []
        for i, (r, tk) in enumerate(zip(reqs, tickets)):
            g = sv.wait(tk, image=True)
            assert g["status"] == 2
            sizes.append(g["batch_size"])
            alone = ctx.render(meshes[r.mesh], [r.view], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)[0][0]
            assert np.array_equal(g["rgba"], alone["rgba"]), i
            x, t, tfs = host[r.mesh]
            if r.mesh not in inputs:
                con, ext, _ = meshes[r.mesh].connectivity()
                xn = O.normalize(x)
                inputs[r.mesh] = dict(xn=xn, t=t, con=con, ext=ext, lmt=O.lmt(xn, t))
            inp = dict(inputs[r.mesh], tf=tfs[tf_ids[r.mesh].index(r.view.tf_id)])
            fp64 = r.mesh not in checked and r.view.width <= 512
            print(f"request {i}: mesh {r.mesh} {r.view.width}^2 against the oracle (fp64 {fp64})", flush=True)
            orc = oracle_ftb(inp, r.view, psi, fp64=fp64)
            assert orc["anomalies"] == 0
            assert np.array_equal(alone["entry_ext"], orc["entry_ext"]), i
            assert np.array_equal(alone["steps"], orc["steps"]), i
            assert np.abs(g["rgba"] - orc["rgba"]).max() <= 2e-3, i
            if fp64:
                cov = orc["entry_ext"] != V.NO_ID
                assert cov.mean() > 0.1
                assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
                checked.add(r.mesh)
        assert max(sizes) > 4
    finally:
        sv.close()
        for m in meshes:
            m.close()
    assert len(checked) == 8


def test_server_pipeline_admits_during_walks(ctx):
    """The pipeline schedule: requests submitted while a long walk is on the
    device (eight 1024^2 views of a 128^3 grid, one walk list) are admitted
    (status 1: preprocessing enqueued, rays in the next walk list) before that
    walk ends, and every image equals the same request rendered alone."""
    import time
    x, t = V.synth_grid(128, 0x5EED0450, 0)
    big_m = ctx.mesh(x, t)
    xs, ts = V.synth_grid(16, 0x5EED0451, 0)
    small_m = ctx.mesh(xs, ts)
    tf = ctx.tf(V.synth_tf(0x5EED0452))
    r = WL.Rng(0x5EED0453, 9)
    big = [V.perspective_view(1024, 1024, tf, eye=2.5 * r.direction()) for _ in range(8)]
    small = [V.perspective_view(96 + 16 * (i % 3), 96 + 16 * (i % 3), tf, eye=2.5 * r.direction()) for i in range(24)]
    print("server pipeline: requests admitted during walks", flush=True)
    sv = V.Server(ctx, [big_m, small_m], hbm_budget=8 << 30, streams=1, max_batch=64, schedule=V.SERVE_PIPELINE)
    try:
        tb = [sv.submit(0, v, keep_image=True) for v in big]
        for _ in range(20000):
            if all(sv.poll(k)["status"] >= 1 for k in tb):
                break
            time.sleep(2e-5)
        mid, seen = [], 0
        for v in small:
            mid.append(sv.submit(1, v, keep_image=True))
            for _ in range(50):
                if sv.poll(mid[-1])["status"] >= 1:
                    # admitted; is a big request admitted before it still walking?
                    seen += any(sv.poll(k)["status"] == 1 for k in tb)
                    break
                time.sleep(2e-5)
        gb = [sv.wait(k, image=True) for k in tb]
        gs = [sv.wait(k, image=True) for k in mid]
    finally:
        sv.close()
    assert seen > 0
    assert max(g["batch_size"] for g in gb + gs) > 1
    for v, g in zip(big, gb):
        alone, _ = ctx.render(big_m, [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        assert g["status"] == 2 and np.array_equal(g["rgba"], alone[0]["rgba"])
    for v, g in zip(small, gs):
        alone, _ = ctx.render(small_m, [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        assert g["status"] == 2 and np.array_equal(g["rgba"], alone[0]["rgba"])
    big_m.close()
    small_m.close()
