"""vfr_render_noalloc (the allocation-free render, through the C ABI) against
vfr_render on the same inputs: entry faces, entry tets, step counts and RGBA
bit for bit, on two calls in a row (the second smaller, reusing the scratch
area the first left behind), in both walk modes; a call beyond the scratch
area fails with VFR_EBUDGET and the next call still renders."""
import ctypes as C

import numpy as np
import pytest

import vfray_amd as V
from vfray_amd import workloads as WL
from helpers import grid_inputs

pytestmark = pytest.mark.gpu


def noalloc(ctx, mesh, views, **kw):
    """Calls vfr_render_noalloc with the options of Context.plan; returns
    (per-view dicts, stats dict)."""
    L = V.lib()
    arr = (V.View * len(views))(*views)
    flags = ((V.RENDER_PERSISTENT if kw.get("persistent") else 0)
             | (V.RENDER_WAVE_QUEUE if kw.get("wave_queue") else 0)
             | (V.RENDER_PSI_TABLE if kw.get("psi_table") else 0))
    o = V.RenderOpts(kw.get("mode", V.MODE_FTB), V.SCHED_BATCHED, kw.get("term_alpha", 0.99), 0, 0, 0,
                     kw.get("max_entries", 0), flags, 0, kw.get("composite", V.COMPOSITE_PPI))
    outs, fos = [], (V.FrameOut * len(views))()
    for i, v in enumerate(views):
        d = {"rgba": np.full((v.height, v.width, 4), np.nan, np.float32),
             "entry_ext": np.full((v.height, v.width), 7, np.uint32),
             "entry_tet": np.full((v.height, v.width), 7, np.uint32),
             "steps": np.full((v.height, v.width), 7, np.uint32),
             "rgb8": np.zeros((v.height, v.width, 3), np.uint8)}
        fos[i].rgba = V.fptr(d["rgba"])
        fos[i].entry_ext = V.uptr(d["entry_ext"])
        fos[i].entry_tet = V.uptr(d["entry_tet"])
        fos[i].steps = V.uptr(d["steps"])
        fos[i].rgb8 = d["rgb8"].ctypes.data_as(C.POINTER(C.c_uint8))
        outs.append(d)
    st = V.Stats()
    rc = L.vfr_render_noalloc(ctx._p, mesh._p, arr, len(views), C.byref(o), fos, C.byref(st))
    if rc != 0:
        raise V.VfrError(rc, L.vfr_last_error().decode())
    return outs, st.as_dict()


def same(a, b):
    for key in ("rgba", "entry_ext", "entry_tet", "steps"):
        assert np.array_equal(a[key], b[key]), key


def test_noalloc_equals_render(ctx):
    inp = grid_inputs(24, 0x5EED0600)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    tf2 = ctx.tf(V.synth_tf(0x5EED0601))
    big = WL.sphere_views(0x5EED0602, 3, [256, 320], tf)
    big.append(V.perspective_view(300, 200, tf2, eye=(2.0, -1.1, 0.8)))
    small = [V.perspective_view(128, 96, tf2, eye=(-1.4, 1.8, 1.0))]
    print("no-alloc render: 4 views, then 1 smaller view", flush=True)
    for views in (big, small):
        got, sg = noalloc(ctx, m, views)
        ref, sr = ctx.render(m, views, mode=V.MODE_FTB)
        assert sg["steps"] == sr["steps"] and sg["active_rays"] == sr["active_rays"] > 0
        for g, r in zip(got, ref):
            same(g, r)
            assert np.array_equal(g["rgb8"], V.to_rgb8(r["rgba"]))
    # the wave-queue walk and the table psi through the same scratch area
    got, _ = noalloc(ctx, m, big[:2], wave_queue=True, psi_table=True)
    ref, _ = ctx.render(m, big[:2], mode=V.MODE_FTB, wave_queue=True, psi_table=True)
    for g, r in zip(got, ref):
        same(g, r)
    # the reference's own mode (orthographic, REF walk), twice, smaller second
    for side in (384, 200):
        v = [V.ortho_view(side, side, tf)]
        got, _ = noalloc(ctx, m, v, mode=V.MODE_REF)
        ref, _ = ctx.render(m, v, mode=V.MODE_REF)
        same(got[0], ref[0])
    m.close()


def test_noalloc_limits(ctx):
    inp = grid_inputs(8, 0x5EED0106)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    print("no-alloc limits", flush=True)
    with pytest.raises(V.VfrError, match="scratch") as e:
        noalloc(ctx, m, [V.perspective_view(4096, 4096, tf, eye=(0, 0, 2.5))])
    assert e.value.code == -6
    with pytest.raises(V.VfrError, match="dropped") as e:
        noalloc(ctx, m, [V.ortho_view(2, 2, tf)])
    assert e.value.code == -8
    v = [V.perspective_view(160, 120, tf, eye=(1.5, 1.9, -0.7))]
    got, _ = noalloc(ctx, m, v)
    ref, _ = ctx.render(m, v, mode=V.MODE_FTB)
    same(got[0], ref[0])
    m.close()
