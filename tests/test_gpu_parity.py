"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Entry faces, entry tets and step counts must be bit-exact; RGBA within
2e-3 max-abs per channel; PSNR >= 50 dB against the FP64 re-integration.
"""
import os

import numpy as np
import pytest

import vfray_amd as V
from vfray_amd import workloads as WL
import oracle as O
from helpers import CONNECTIVITY_CASES, brute_con, c1_inputs, connectivity_case, grid_inputs, oracle_ftb, psnr

pytestmark = pytest.mark.gpu

TOL = 2e-3


@pytest.fixture(scope="module")
def psi():
    return V.psi_generate()


@pytest.fixture(scope="module")
def c1():
    return c1_inputs()


def _ext_tet(ext, entry):
    out = np.full(entry.shape, V.NO_ID, np.uint32)
    m = entry != V.NO_ID
    out[m] = ext[entry[m], 0]
    return out


@pytest.mark.parametrize("morton", [False, True])
def test_connectivity_matches_oracle(ctx, c1, morton):
    print(f"connectivity 16^3 morton={morton}", flush=True)
    m = ctx.mesh(c1["x"], c1["t"], morton=morton)
    con, ext, lmt = m.connectivity()
    assert np.array_equal(con, c1["con"])
    assert np.array_equal(ext, c1["ext"])
    assert np.array_equal(lmt.view(np.uint32), c1["lmt"].view(np.uint32))
    assert np.array_equal(m.vertices().view(np.uint32), c1["xn"].view(np.uint32))
    m.close()


def test_connectivity_given_files(ctx, c1):
    """con/extf/lmt passed in (the .con/.extf/.lmt cache path) are used as is."""
    m = ctx.mesh(c1["x"], c1["t"], con=c1["con"], extf=c1["ext"], lmt=c1["lmt"])
    con, ext, lmt = m.connectivity()
    assert np.array_equal(con, c1["con"]) and np.array_equal(ext, c1["ext"])
    m.close()


def test_mesh_checks_and_limits(ctx):
    """Mesh creation checks every id on the device (first offender named, as
    the host checks did), and builds buildLmt's limits on the device bit for
    bit, including its first-wins rule where the z extreme is a signed zero."""
    x = np.array([[0, 0, -1.5, 0.1], [1, 0, -0.0, 0.2], [0, 1, 0.0, 0.3], [0, 0, -2, 0.4]], np.float32)
    t = np.array([[0, 1, 2, 3]], np.uint32)
    for z in ([-1.5, -0.0, 0.0, -2.0], [0.5, 0.0, -0.0, 2.0], [0.25, 0.5, 0.75, 0.125]):
        xz = x.copy()
        xz[:, 2] = np.array(z, np.float32)
        m = ctx.mesh(xz, t, normalize=False)
        lmt = m.connectivity()[2]
        assert np.array_equal(lmt.view(np.uint32), O.lmt(xz, t).view(np.uint32)), (z, lmt)
        m.close()
    t2 = np.array([[0, 1, 2, 3], [0, 1, 2, 9]], np.uint32)
    with pytest.raises(V.VfrError, match="tet 1 references vertex 9 >= 4"):
        ctx.mesh(x, t2)
    with pytest.raises(V.VfrError, match="repeated vertex"):
        ctx.mesh(x, np.array([[0, 1, 2, 2]], np.uint32))
    con = np.array([[0, 0, 0, 5]], np.uint32)
    with pytest.raises(V.VfrError, match="con entry 3 = 5 >= 1"):
        ctx.mesh(x, t, con=con)
    con = np.zeros((1, 4), np.uint32)
    with pytest.raises(V.VfrError, match="extf entry 2 invalid"):
        ctx.mesh(x, t, con=con, extf=np.array([[0, 0], [0, 1], [0, 4], [0, 3]], np.uint32))


def test_connectivity_32(ctx):
    inp = grid_inputs(32, 0x5EED0100)
    print("connectivity 32^3", flush=True)
    m = ctx.mesh(inp["x"], inp["t"])
    con, ext, lmt = m.connectivity()
    assert np.array_equal(con, inp["con"])
    assert np.array_equal(ext, inp["ext"])
    assert np.array_equal(lmt.view(np.uint32), inp["lmt"].view(np.uint32))
    m.close()


@pytest.mark.parametrize("case", CONNECTIVITY_CASES)
@pytest.mark.parametrize("morton", [False, True])
def test_connectivity_brute_force(ctx, case, morton):
    """The device's sort-and-match connectivity against exhaustive search,
    including a face shared by three tets (the smallest other tet wins) and
    single-tet and 1^3..3^3 grids (every face external on one tet)."""
    print(f"connectivity {case} morton={morton}", flush=True)
    x, t = connectivity_case(case)
    m = ctx.mesh(x, t, morton=morton)
    con, ext, _ = m.connectivity()
    bc = brute_con(t)
    assert np.array_equal(con, bc)
    own = np.arange(t.shape[0])[:, None]
    want_ext = np.argwhere(bc == own).astype(np.uint32)   # (tet, face), tet-major order
    assert np.array_equal(ext, want_ext)
    assert m.info["n_ext_faces"] == want_ext.shape[0]
    m.close()


def test_ref_mode_c1(ctx, c1, psi):
    """512^2 orthographic REF render: the reference's semantics end to end."""
    print("REF 16^3 512x512", flush=True)
    orc = O.render_ref(c1["xn"], c1["t"], c1["con"], c1["ext"], c1["tf"], psi, c1["lmt"], 512, 512)
    assert orc["face_returns"] == 0
    m = ctx.mesh(c1["x"], c1["t"])
    tf = ctx.tf(c1["tf"])
    outs, st = ctx.render(m, [V.ortho_view(512, 512, tf)], mode=V.MODE_REF)
    g = outs[0]
    assert np.array_equal(g["entry_ext"], orc["entry_ext"])
    assert np.array_equal(g["entry_tet"], _ext_tet(c1["ext"], orc["entry_ext"]))
    assert np.array_equal(g["steps"], orc["steps"])
    assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL
    assert np.abs(V.to_rgb8(g["rgba"]).astype(int) - orc["rgb8"]).max() <= 1
    assert st["steps"] == int(orc["steps"].sum())
    assert st["active_rays"] == int((orc["entry_ext"] != V.NO_ID).sum())
    m.close()


def _rot(a, b):
    """A 3x4 orthographic view matrix: rotation about z by a, then about x by b (radians)."""
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])
    return np.hstack([rx @ rz, np.zeros((3, 1))]).astype(np.float32)


def test_ref_face_memo(ctx, c1, psi):
    """The face-plane memo (VFR_RENDER_FACE_MEMO) on the REF walk: the
    reference's own case renders bit for bit as the walk that solves each
    exit face (entry ids, steps, RGBA) and as the oracle, per-view and
    batched, in PPI and linear compositing; a batch of rotated views (one
    memo table per view) equals the solving walk view by view; the
    autotuner's REF choice renders the same bits."""
    print("REF face memo 16^3 512x512", flush=True)
    m = ctx.mesh(c1["x"], c1["t"])
    tf = ctx.tf(c1["tf"])
    for composite in (V.COMPOSITE_PPI, V.COMPOSITE_LINEAR_OPACITY):
        orc = O.render_ref(c1["xn"], c1["t"], c1["con"], c1["ext"], c1["tf"], psi, c1["lmt"], 512, 512,
                           composite=composite)
        for sched in (V.SCHED_PER_VIEW, V.SCHED_BATCHED):
            a, sa = ctx.render(m, [V.ortho_view(512, 512, tf)], mode=V.MODE_REF, schedule=sched,
                               composite=composite)
            b, sb = ctx.render(m, [V.ortho_view(512, 512, tf)], mode=V.MODE_REF, schedule=sched,
                               composite=composite, face_memo=True)
            for key in ("rgba", "entry_ext", "steps"):
                assert np.array_equal(a[0][key], b[0][key]), (composite, sched, key)
            assert np.array_equal(b[0]["entry_ext"], orc["entry_ext"])
            assert np.array_equal(b[0]["steps"], orc["steps"])
            assert np.abs(b[0]["rgba"] - orc["rgba"]).max() <= TOL
            assert sa["steps"] == sb["steps"] == int(orc["steps"].sum())
    print("REF face memo, 4 rotated views", flush=True)
    views = [V.ortho_view(160 + 16 * i, 144, tf, _rot(0.4 * i + 0.1, 0.3 * i - 0.5)) for i in range(4)]
    a, sa = ctx.render(m, views, mode=V.MODE_REF, schedule=V.SCHED_BATCHED)
    b, sb = ctx.render(m, views, mode=V.MODE_REF, schedule=V.SCHED_BATCHED, face_memo=True)
    assert sa["steps"] == sb["steps"] and sa["active_rays"] > 0
    for ga, gb in zip(a, b):
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(ga[key], gb[key]), key
    p = ctx.plan(m, views, mode=V.MODE_REF, schedule=V.SCHED_BATCHED)
    choice = p.autotune()
    assert isinstance(choice["face_memo"], bool) and choice["traversal_ms"] > 0
    p.execute()
    for i in range(len(views)):
        assert np.array_equal(p.read(i)["rgba"], a[i]["rgba"])
    p.close()
    m.close()


@pytest.mark.parametrize("composite", [V.COMPOSITE_LINEAR, V.COMPOSITE_LINEAR_OPACITY])
def test_ref_linear_c1(ctx, c1, psi, composite):
    """Linear compositing (USEPARTIALPREINT 0) on the REF walk: image bytes,
    RGBA, entry faces and step counts bit-exact with the oracle, per-view and
    batched alike."""
    print(f"REF linear composite={composite} 16^3 512x512", flush=True)
    orc = O.render_ref(c1["xn"], c1["t"], c1["con"], c1["ext"], c1["tf"], psi, c1["lmt"], 512, 512,
                       composite=composite)
    m = ctx.mesh(c1["x"], c1["t"])
    tf = ctx.tf(c1["tf"])
    for sched in (V.SCHED_PER_VIEW, V.SCHED_BATCHED):
        outs, st = ctx.render(m, [V.ortho_view(512, 512, tf)], mode=V.MODE_REF, schedule=sched,
                              composite=composite)
        g = outs[0]
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.array_equal(g["rgba"], orc["rgba"])
        assert np.array_equal(V.to_rgb8(g["rgba"]), orc["rgb8"])
        assert st["steps"] == int(orc["steps"].sum())
    assert orc["rgb8"].max() > 0
    m.close()


@pytest.mark.parametrize("composite", [V.COMPOSITE_LINEAR, V.COMPOSITE_LINEAR_OPACITY])
def test_ftb_linear_equals_oracle(ctx, psi, composite):
    """Linear compositing on the front-to-back walk: perspective views of the
    voided, graded mesh (rays exit and re-enter), batched and per-view, bit
    for bit with the oracle (entry faces, steps, RGBA); the walk-kernel flags
    for PPI mode do not change the result."""
    inp = grid_inputs(20, 0x5EED0303, variant=1)
    tfa = inp["tf"].copy().reshape(256, 4)
    tfa[:, 3] *= 0.5
    inp = dict(inp, tf=tfa.reshape(-1))
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    from vfray_amd import workloads as WL
    views = WL.sphere_views(0x5EED0304, 3, [192, 256], tf)
    print(f"FTB linear composite={composite} concave 20^3", flush=True)
    a, sa = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, composite=composite)
    b, _ = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW, composite=composite)
    c, _ = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, composite=composite,
                      persistent=True)
    assert sa["anomalies"] == 0
    multi = 0
    for v, g, h, k in zip(views, a, b, c):
        orc = oracle_ftb(inp, v, psi, composite=composite)
        multi += int((orc["n_entries"] > 1).sum())
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(g[key], h[key]) and np.array_equal(g[key], k[key]), key
            assert np.array_equal(g[key], orc[key]), key
        # not a black image: most covered pixels get colour (the oracle's
        # render of these views: above 82 %), and with opacity some alpha
        cov = orc["entry_ext"] != V.NO_ID
        assert (g["rgba"][cov][:, :3].max(1) > 0).mean() > 0.5
        if composite == V.COMPOSITE_LINEAR_OPACITY:
            assert g["rgba"][..., 3].max() > 0.1
    assert multi > 500
    m.close()
    m.close()


def test_mesh_rotate_matches_reference(ctx, psi):
    """vfr_mesh_rotate (the viewer's rotate + renormalize, on the device):
    two presses at (30, 20) on the 8^3 mesh give the compiled reference's
    vertices bit for bit; the REF render with the kept (stale) limits has its
    owner faces and equals the oracle's image; a plan made before the
    rotation refuses to run."""
    rot = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_rot.npz")))
    x, t, tf, _ = WL.plain_cpu(8)
    print("mesh rotate 8^3 x2", flush=True)
    m = ctx.mesh(x, t)
    tfi = ctx.tf(tf)
    p = ctx.plan(m, [V.ortho_view(128, 128, tfi)], mode=V.MODE_REF)
    lmt0 = m.connectivity()[2]
    for mv in rot["mv"]:
        m.rotate(mv)
    assert np.array_equal(m.vertices().view(np.uint32), rot["verts"].view(np.uint32))
    assert np.array_equal(m.connectivity()[2], lmt0)
    with pytest.raises(V.VfrError, match="rotated"):
        p.execute()
    p.close()
    outs, _ = ctx.render(m, [V.ortho_view(128, 128, tfi)], mode=V.MODE_REF)
    g = outs[0]
    assert np.array_equal(g["entry_ext"], rot["owner128"])
    xn = O.normalize(x)
    con, _ = O.connectivity(t, x.shape[0])
    orc = O.render_ref(rot["verts"], t, con, O.extf(con), tf, psi, O.lmt(xn, t), 128, 128)
    assert np.array_equal(g["rgba"], orc["rgba"])
    assert np.array_equal(g["steps"], orc["steps"])
    m.close()


def test_ftb_psi_table(ctx, c1, psi):
    """VFR_RENDER_PSI_TABLE: front-to-back with psi looked up in the table, bit
    for bit with the oracle's table variant, for the one-ray-per-thread and
    the persistent kernels; the series default stays within the table's
    quantization of it and closer to the FP64 re-integration.  The
    64-bit-address walk variant (VFR_RENDER_WIDE) gives the same bits."""
    print("FTB psi table 16^3", flush=True)
    m = ctx.mesh(c1["x"], c1["t"])
    tf = ctx.tf(c1["tf"])
    views = [V.ortho_view(512, 512, tf), V.ortho_view(300, 200, tf)]
    outs, _ = ctx.render(m, views, mode=V.MODE_FTB, psi_table=True)
    outp, _ = ctx.render(m, views, mode=V.MODE_FTB, psi_table=True, persistent=True)
    outs_s, _ = ctx.render(m, views, mode=V.MODE_FTB)
    for v, g, gp, gs in zip(views, outs, outp, outs_s):
        orc = oracle_ftb(c1, v, psi, fp64=True, psi_table=True)
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.array_equal(g["rgba"], orc["rgba"])
        assert np.array_equal(gp["rgba"], g["rgba"]) and np.array_equal(gp["steps"], g["steps"])
        cov = orc["entry_ext"] != V.NO_ID
        assert np.abs(gs["rgba"] - g["rgba"]).max() <= TOL
        assert psnr(gs["rgba"], orc["rgba64"], cov) > psnr(g["rgba"], orc["rgba64"], cov)
    # the 64-bit-address variant (records of 2^27 tets or more) on this mesh
    outw, _ = ctx.render(m, views, mode=V.MODE_FTB, wide=True)
    outwt, _ = ctx.render(m, views, mode=V.MODE_FTB, wide=True, psi_table=True)
    for g, gw, gs, gwt in zip(outs, outw, outs_s, outwt):
        assert np.array_equal(gw["rgba"], gs["rgba"]) and np.array_equal(gw["steps"], gs["steps"])
        assert np.array_equal(gwt["rgba"], g["rgba"]) and np.array_equal(gwt["steps"], g["steps"])
    m.close()


@pytest.mark.parametrize("schedule", [V.SCHED_PER_VIEW, V.SCHED_BATCHED])
def test_ftb_ortho_c1(ctx, c1, psi, schedule):
    print(f"FTB ortho 16^3 schedule={schedule}", flush=True)
    m = ctx.mesh(c1["x"], c1["t"])
    tf = ctx.tf(c1["tf"])
    views = [V.ortho_view(512, 512, tf), V.ortho_view(300, 200, tf)]
    outs, st = ctx.render(m, views, mode=V.MODE_FTB, schedule=schedule)
    for v, g in zip(views, outs):
        orc = oracle_ftb(c1, v, psi, fp64=True)
        assert orc["anomalies"] == 0
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL
        cov = orc["entry_ext"] != V.NO_ID
        assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m.close()


def _sphere_views(tf, sides, seed=0x5EED0101):
    from vfray_amd import workloads as WL
    return WL.sphere_views(seed, len(sides), list(sides), tf)


def test_ftb_perspective_32(ctx, psi):
    inp = grid_inputs(32, 0x5EED0102)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    from vfray_amd import workloads as WL
    views = WL.sphere_views(0x5EED0102, 3, [384], tf)
    print("FTB perspective 32^3 3x384^2", flush=True)
    outs, st = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    for v, g in zip(views, outs):
        orc = oracle_ftb(inp, v, psi, fp64=True)
        assert orc["anomalies"] == 0
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL
        cov = orc["entry_ext"] != V.NO_ID
        assert cov.sum() > 0.2 * cov.size
        assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m.close()


def test_ftb_rgba_bit_exact(ctx, psi):
    """The front-to-back walk's colours equal the oracle's bit for bit, not
    only within the tolerance: the kernel and the oracle perform the same
    float operations in the same order (fused multiply-adds written out,
    no contraction), so any change to the kernel's arithmetic that is not
    mirrored in the oracle shows here.  A transfer function whose
    neighbouring entries differ makes psi count in every segment; a concave
    mesh adds re-entries; perspective and orthographic views."""
    rng = np.random.default_rng(0x5EED010A)
    tf = np.empty((256, 4), np.float32)
    tf[:, :3] = rng.random((256, 3), dtype=np.float32)
    tf[:, 3] = 0.6 * rng.random(256, dtype=np.float32)
    for n, variant, seed in ((16, 0, 0x5EED010A), (12, 1, 0x5EED010B)):
        inp = grid_inputs(n, seed, variant)
        inp["tf"] = tf
        m = ctx.mesh(inp["x"], inp["t"])
        tid = ctx.tf(tf)
        views = WL.sphere_views(seed, 3, [64, 96], tid)
        views.append(V.ortho_view(80, 72, tid, matrix=_rot(0.3, -0.4)))
        print(f"FTB bit-exact colours {n}^3 variant {variant}", flush=True)
        outs, _ = ctx.render(m, views, mode=V.MODE_FTB)
        for g, v in zip(outs, views):
            orc = oracle_ftb(inp, v, psi)
            assert (orc["entry_ext"] != V.NO_ID).sum() > 0
            assert np.array_equal(g["entry_ext"], orc["entry_ext"])
            assert np.array_equal(g["steps"], orc["steps"])
            assert np.array_equal(g["rgba"].view(np.uint32), orc["rgba"].view(np.uint32))
        m.close()


def test_batched_thin_views(ctx, psi):
    """A batch of views whose 8x8 tiles are partly filled (2 pixels wide or
    tall, odd sides): the batched plan builds its ray list per tile, the
    per-view schedule in one pass; both equal each other and the oracle."""
    inp = grid_inputs(12, 0x5EED0109)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    r = WL.Rng(0x5EED0109, 9)
    sizes = [(2, 64), (64, 2), (9, 9), (7, 300), (2, 2), (130, 3), (17, 33), (3, 129)]
    views = [V.perspective_view(w, h, tf, eye=2.5 * r.direction()) for w, h in sizes]
    print("thin views batched vs per-view vs oracle 12^3", flush=True)
    a, sa = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    b, sb = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
    c, sc = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, pipelined=True)
    assert sa["steps"] == sb["steps"] == sc["steps"] and sa["active_rays"] == sb["active_rays"] > 0
    for i, v in enumerate(views):
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(a[i][key], b[i][key])
            assert np.array_equal(a[i][key], c[i][key])
    for i in (2, 6):
        orc = oracle_ftb(inp, views[i], psi)
        assert np.array_equal(a[i]["entry_ext"], orc["entry_ext"])
        assert np.array_equal(a[i]["steps"], orc["steps"])
        assert np.abs(a[i]["rgba"] - orc["rgba"]).max() <= TOL
    m.close()


def test_raster_face_sizes(ctx, psi):
    """Entry lists from faces of every size the raster splits its work by: a
    3^3 grid at 512^2 (bounding boxes of thousands of pixels, rasterized by
    the whole wave in 8x8 tiles), a 6^3 grid at 160^2 (a few hundred pixels
    each, shared out over several wave iterations) and views of a few pixels
    per face (16 and 64 faces per wave), in one launch and one per view,
    against the oracle bit for bit."""
    for n, seed, sides in ((3, 0x5EED010C, (512, 40)), (6, 0x5EED010D, (160, 24))):
        inp = grid_inputs(n, seed)
        m = ctx.mesh(inp["x"], inp["t"])
        tf = ctx.tf(inp["tf"])
        views = [WL.sphere_views(seed + k, 1, [s], tf)[0] for k, s in enumerate(sides)]
        views.append(V.ortho_view(sides[0], sides[0] - 37, tf, matrix=_rot(0.7, 0.2)))
        print(f"raster face sizes {n}^3 {sides}", flush=True)
        a, _ = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
        b, _ = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        for i, v in enumerate(views):
            orc = oracle_ftb(inp, v, psi)
            assert (orc["entry_ext"] != V.NO_ID).sum() > 0
            for g in (a[i], b[i]):
                assert np.array_equal(g["entry_ext"], orc["entry_ext"])
                assert np.array_equal(g["steps"], orc["steps"])
                assert np.array_equal(g["rgba"].view(np.uint32), orc["rgba"].view(np.uint32))
        m.close()


def test_batched_equals_per_view(ctx):
    """Each view rendered in a batch equals the same view rendered alone, with
    every walk kernel (one ray per thread with and without LDS-staged tet
    records, lane-refill persistent, wave work queue) and with pipelined view
    chunks."""
    inp = grid_inputs(24, 0x5EED0103)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    from vfray_amd import workloads as WL
    views = WL.sphere_views(0x5EED0103, 5, [256, 384, 512], tf)
    print("batched vs per-view 24^3", flush=True)
    a, sa = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    b, sb = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
    c, sc = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, persistent=True)
    d, sd = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, persistent=True, waves_per_cu=4,
                       refill=1)
    e, se = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, pipelined=True)
    f, sf = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, wave_queue=True)
    g, sg = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, wave_queue=True, pipelined=True)
    h, sh = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, lds_records=True)
    k, sk = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW, lds_records=True)
    assert sa["steps"] == sb["steps"] == sc["steps"] == sd["steps"] == se["steps"] == sf["steps"] == sg["steps"]
    assert sa["steps"] == sh["steps"] == sk["steps"]
    assert sa["active_rays"] == sb["active_rays"] == se["active_rays"]
    for i, v in enumerate(views):
        alone, _ = ctx.render(m, [v], mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(a[i][key], b[i][key])
            assert np.array_equal(a[i][key], c[i][key])
            assert np.array_equal(a[i][key], d[i][key])
            assert np.array_equal(a[i][key], e[i][key])
            assert np.array_equal(a[i][key], f[i][key])
            assert np.array_equal(a[i][key], g[i][key])
            assert np.array_equal(a[i][key], h[i][key])
            assert np.array_equal(a[i][key], k[i][key])
            assert np.array_equal(a[i][key], alone[0][key])
    m.close()


def test_graph_replays(ctx):
    """A captured batched plan replayed several times, with eager renders in
    between, gives the eager result every time."""
    inp = grid_inputs(20, 0x5EED0104)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    from vfray_amd import workloads as WL
    views = WL.sphere_views(0x5EED0104, 3, [256, 320], tf)
    print("graph replays 20^3", flush=True)
    ref, _ = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    for wq in (False, True):
        p = ctx.plan(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, use_graph=True, wave_queue=wq)
        for rep in range(3):
            p.execute()
            for i in range(len(views)):
                got = p.read(i)
                for key in ("rgba", "entry_ext", "steps"):
                    assert np.array_equal(got[key], ref[i][key]), (wq, rep, i, key)
            ctx.render(m, views[:1], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        p.close()
    m.close()


def test_ftb_c2_full(ctx, psi):
    """64^3 grid (1.57M tets), one 1024^2 perspective view, alpha 0.99."""
    from vfray_amd import workloads as WL
    x, t, tfa, seed = WL.plain_gpu()
    inp = dict(x=x, t=t, tf=tfa, xn=O.normalize(x))
    inp["con"], _ = O.connectivity(t, x.shape[0])
    inp["ext"] = O.extf(inp["con"])
    inp["lmt"] = O.lmt(inp["xn"], t)
    m = ctx.mesh(x, t)
    tf = ctx.tf(tfa)
    v = WL.sphere_views(seed, 1, [1024], tf)[0]
    print("FTB 64^3 1024x1024", flush=True)
    outs, st = ctx.render(m, [v], mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
    g = outs[0]
    orc = oracle_ftb(inp, v, psi, fp64=True)
    assert np.array_equal(g["entry_ext"], orc["entry_ext"])
    assert np.array_equal(g["steps"], orc["steps"])
    assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL
    cov = orc["entry_ext"] != V.NO_ID
    assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m.close()


def test_block_tf_staging(ctx):
    """A batched plan whose views use more transfer functions than the walk
    can stage in LDS for the launch: each 256-ray block stages the TFs of its
    own rays' views, and blocks whose rays span more than four views (here
    the run of 8x8 views) read the global table.  Every view renders bit
    for bit as it does alone (one TF, staged for the launch)."""
    inp = grid_inputs(14, 0x5EED0107)
    m = ctx.mesh(inp["x"], inp["t"])
    tfs = [ctx.tf(V.synth_tf(0x5EED0107 + k)) for k in range(7)]
    from vfray_amd import workloads as WL
    views = []
    for k in range(14):
        side = 8 if 4 <= k < 11 else 96 + 16 * (k % 3)
        views += WL.sphere_views(0x5EED0107 + k, 1, [side], tfs[k % len(tfs)])
    print("per-block TF staging, 14 views, 7 TFs", flush=True)
    a, sa = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    assert sa["anomalies"] == 0 and sa["active_rays"] > 0
    for i, v in enumerate(views):
        b, _ = ctx.render(m, [v], mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(a[i][key], b[0][key]), (i, key)
    m.close()


def test_autotune_keeps_results(ctx):
    """The autotuner's chosen schedule renders the same bits."""
    inp = grid_inputs(16, 0x5EED0105)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    from vfray_amd import workloads as WL
    views = WL.sphere_views(0x5EED0105, 4, [128, 192], tf)
    print("autotune 16^3", flush=True)
    ref, _ = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
    p = ctx.plan(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    choice = p.autotune()
    assert 1 <= choice["waves_per_cu"] <= 64 and 1 <= choice["refill"] <= 64
    assert isinstance(choice["persistent"], bool)
    p.execute()
    for i in range(len(views)):
        assert np.array_equal(p.read(i)["rgba"], ref[i]["rgba"])
    p.close()
    m.close()


@pytest.mark.parametrize("term", [0.99, 1.0])
def test_ftb_concave_reentry(ctx, psi, term):
    """Voided, radially graded mesh: rays exit through interior boundaries and
    re-enter; batched and per-view renders equal the oracle bit for bit, and
    so do the other walk kernels (lane-refill persistent, whose idle lanes
    also step and may try a re-entry; wave work queue; LDS-staged records)
    and the pipelined schedule."""
    inp = grid_inputs(20, 0x5EED0303, variant=1)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    from vfray_amd import workloads as WL
    views = WL.sphere_views(0x5EED0303, 4, [256, 320], tf)
    print(f"FTB concave 20^3 term={term}", flush=True)
    a, sa = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, term_alpha=term)
    b, sb = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW, term_alpha=term)
    assert sa["anomalies"] == 0
    for kw in (dict(persistent=True), dict(persistent=True, waves_per_cu=4, refill=1), dict(wave_queue=True),
               dict(lds_records=True), dict(pipelined=True)):
        c, sc = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, term_alpha=term, **kw)
        assert sc["steps"] == sa["steps"] and sc["anomalies"] == 0, kw
        for g, h in zip(a, c):
            for key in ("rgba", "entry_ext", "steps"):
                assert np.array_equal(g[key], h[key]), (kw, key)
    multi = 0
    for v, g, h in zip(views, a, b):
        orc = oracle_ftb(inp, v, psi, term=term, fp64=True)
        assert orc["anomalies"] == 0
        multi += int((orc["n_entries"] > 1).sum())
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(g[key], h[key]), key
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL
        cov = orc["entry_ext"] != V.NO_ID
        assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    assert multi > 1000  # pixels whose ray meets the boundary more than once
    m.close()


def test_ftb_concave_c4_views(ctx, psi):
    """The concave benchmark mesh (96^3 grid, graded, 12 voids), all 32 views
    at 768^2 against the oracle (entry ids and steps bit for bit, RGBA within
    2e-3), two of them also against its FP64 re-integration (PSNR >= 50 dB);
    too few entry slots are reported, more render the same."""
    from vfray_amd import workloads as WL
    x, t, tfa, seed = WL.concave()
    inp = dict(x=x, t=t, tf=tfa, xn=O.normalize(x))
    inp["con"], _ = O.connectivity(t, x.shape[0])
    inp["ext"] = O.extf(inp["con"])
    inp["lmt"] = O.lmt(inp["xn"], t)
    m = ctx.mesh(x, t)
    tf = ctx.tf(tfa)
    views = WL.sphere_views(seed, 32, [768], tf)
    print("FTB concave 96^3 32x768^2", flush=True)
    # up to 13 entry faces per pixel here: 8 slots drop some, and say so
    with pytest.raises(V.VfrError, match="dropped"):
        ctx.render(m, views[:2], mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, max_entries=8)
    outs, st = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    assert st["anomalies"] == 0
    # 32 slot planes instead of the 16 chosen: the same lists, the same images
    outs32, _ = ctx.render(m, views[:2], mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, max_entries=32)
    for g, h in zip(outs, outs32):
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(g[key], h[key]), key
    for i, (v, g) in enumerate(zip(views, outs)):
        print(f"oracle concave view {i}", flush=True)
        orc = oracle_ftb(inp, v, psi, fp64=i < 2)
        assert orc["anomalies"] == 0
        assert np.array_equal(g["entry_ext"], orc["entry_ext"]), i
        assert np.array_equal(g["steps"], orc["steps"]), i
        assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL, i
        if i < 2:
            cov = orc["entry_ext"] != V.NO_ID
            assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m.close()


def test_ftb_batch64_full(ctx, psi):
    """The benchmark's workload at full size: 128^3 grid (12.58 M tets), all
    64 views.  Batched equals per-view bit for bit on every view, the stats
    add up, no anomalies; every view equals the oracle (entry ids and steps
    bit for bit, RGBA within 2e-3), and three of them (256^2, 384^2, 512^2)
    also meet PSNR >= 50 dB against its FP64 re-integration."""
    from vfray_amd import workloads as WL
    x, t, tfa, seed = WL.batch64()
    m = ctx.mesh(x, t)
    tf = ctx.tf(tfa)
    views = WL.sphere_views(seed, 64, list(WL.SIDES_BATCH), tf)
    print("FTB 128^3 64 views batched vs per-view", flush=True)
    a, sa = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
    b, sb = ctx.render(m, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
    assert sa["anomalies"] == 0 and sa["steps"] == sb["steps"]
    assert sa["steps"] == sum(int(g["steps"].sum()) for g in a)
    assert sa["active_rays"] == sum(int((g["entry_ext"] != V.NO_ID).sum()) for g in a)
    for ga, gb in zip(a, b):
        for key in ("rgba", "entry_ext", "steps"):
            assert np.array_equal(ga[key], gb[key])
    inp = dict(x=x, t=t, tf=tfa, xn=O.normalize(x))
    inp["con"], _ = O.connectivity(t, x.shape[0])
    inp["ext"] = O.extf(inp["con"])
    inp["lmt"] = O.lmt(inp["xn"], t)
    fp64_views = {[k for k, v in enumerate(views) if v.width == side][0] for side in (256, 384, 512)}
    for i, v in enumerate(views):
        print(f"oracle view {i} {v.width}^2", flush=True)
        orc = oracle_ftb(inp, v, psi, fp64=i in fp64_views)
        g = a[i]
        assert orc["anomalies"] == 0
        assert np.array_equal(g["entry_ext"], orc["entry_ext"]), i
        assert np.array_equal(g["steps"], orc["steps"]), i
        assert np.abs(g["rgba"] - orc["rgba"]).max() <= TOL, i
        cov = orc["entry_ext"] != V.NO_ID
        assert cov.mean() > 0.2
        if i in fp64_views:
            assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m.close()


def test_edge_cases(ctx, psi):
    """A camera whose field of view misses the mesh (empty image, no rays),
    one with the mesh behind it (refused: perspective views need the whole
    mesh in front of the eye), the smallest image (2x2, both modes), a ray
    along the mesh's boundary edge (more entry faces than a list holds:
    reported), and a single tet seen in perspective, against the oracle."""
    inp = grid_inputs(8, 0x5EED0106)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    print("edge cases", flush=True)
    off = V.perspective_view(64, 48, tf, eye=(0, 0, 5.0), target=(3.0, 0, 0), fov_y_deg=20.0)
    outs, st = ctx.render(m, [off], mode=V.MODE_FTB)
    assert st["active_rays"] == 0 and st["steps"] == 0
    assert not outs[0]["rgba"].any() and (outs[0]["entry_ext"] == V.NO_ID).all()
    behind = V.perspective_view(64, 48, tf, eye=(0, 0, 3.0), target=(0, 0, 6.0))
    with pytest.raises(V.VfrError, match="in front|inside or beside"):
        ctx.render(m, [behind], mode=V.MODE_FTB)
    for mode in (V.MODE_FTB, V.MODE_REF):
        # front-to-back: scaled so that the four pixel rays do not run along
        # the mesh's boundary edges (a ray along an edge enters every face
        # on it, more than any entry list holds: reported as an error)
        mat = np.diag([1.2, 1.2, 1.2]).astype(np.float32) if mode == V.MODE_FTB else np.eye(3, dtype=np.float32)
        tiny = V.ortho_view(2, 2, tf, matrix=np.hstack([mat, np.zeros((3, 1), np.float32)]))
        g = ctx.render(m, [tiny], mode=mode)[0][0]
        if mode == V.MODE_FTB:
            orc = oracle_ftb(inp, tiny, psi)
        else:
            orc = O.render_ref(inp["xn"], inp["t"], inp["con"], inp["ext"], inp["tf"], psi, inp["lmt"], 2, 2)
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.array_equal(g["rgba"], orc["rgba"])
    with pytest.raises(V.VfrError, match="dropped"):
        ctx.render(m, [V.ortho_view(2, 2, tf)], mode=V.MODE_FTB)
    m.close()
    x = np.array([[0, 0, 0, 0.1], [1, 0, 0, 0.9], [0, 1, 0, 0.5], [0.3, 0.3, 1, 0.3]], np.float32)
    tt = np.array([[0, 1, 2, 3]], np.uint32)
    one = dict(x=x, t=tt, tf=inp["tf"], xn=O.normalize(x))
    one["con"], _ = O.connectivity(tt, 4)
    one["ext"] = O.extf(one["con"])
    one["lmt"] = O.lmt(one["xn"], tt)
    m1 = ctx.mesh(x, tt)
    v = V.perspective_view(96, 96, tf, eye=(1.5, -2.0, 1.2))
    g = ctx.render(m1, [v], mode=V.MODE_FTB)[0][0]
    orc = oracle_ftb(one, v, psi, fp64=True)
    cov = orc["entry_ext"] != V.NO_ID
    assert cov.sum() > 200
    assert np.array_equal(g["entry_ext"], orc["entry_ext"])
    assert np.array_equal(g["steps"], orc["steps"])
    assert np.array_equal(g["rgba"], orc["rgba"])
    assert (g["steps"][cov] == 1).all()
    assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m1.close()


def test_dropped_entries_reported_without_stats(ctx):
    """Rays along a grid's boundary edges (a 2x2 orthographic view: 129 entry
    faces at a pixel): a plan with 8 slots reports VFR_EMESH from
    vfr_plan_execute whether or not the caller asks for stats, and a plan
    left to choose its slots refuses to be created (more than 32)."""
    import ctypes as C
    inp = grid_inputs(8, 0x5EED0106)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    print("dropped entries, stats NULL", flush=True)
    p = ctx.plan(m, [V.ortho_view(2, 2, tf)], mode=V.MODE_FTB, max_entries=8)
    L = V.lib()
    assert L.vfr_plan_execute(p._p, None) == -8
    assert "dropped" in L.vfr_last_error().decode()
    st = V.Stats()
    assert L.vfr_plan_execute(p._p, C.byref(st)) == -8
    p.close()
    with pytest.raises(V.VfrError, match="more than the 32 entry slots") as e:
        ctx.plan(m, [V.ortho_view(2, 2, tf)], mode=V.MODE_FTB)
    assert e.value.code == -8
    # a view whose rays stay off the boundary (scaled by 1.2: at most one entry
    # face per pixel) renders on the same mesh with 8 slots
    mat = np.hstack([np.diag([1.2, 1.2, 1.2]), np.zeros((3, 1))]).astype(np.float32)
    ok = ctx.plan(m, [V.ortho_view(64, 48, tf, matrix=mat)], mode=V.MODE_FTB, max_entries=8)
    assert L.vfr_plan_execute(ok._p, None) == 0
    ok.close()
    m.close()
