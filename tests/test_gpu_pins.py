"""The closed-form pins of test_cpu_pins.py on the HIP path (through the C
ABI): the constant-TF chord opacity and colour on square, non-square and
orthographic views, and step counts equal to the tets each pixel's line
crosses, batched and per-view."""
import numpy as np
import pytest

import vfray_amd as V
import oracle as O
from test_cpu_pins import (ALPHA, C_RGB, _grid, check_chord, check_steps, chord_views, step_cases)
from closed_forms import constant_tf, cube_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("schedule", [V.SCHED_BATCHED, V.SCHED_PER_VIEW])
def test_gpu_chord_closed_form(ctx, schedule):
    print(f"chord closed form, schedule {schedule}", flush=True)
    n = 8
    x, t = cube_grid(n)
    m = ctx.mesh(x, t)
    tf = ctx.tf(constant_tf(C_RGB, ALPHA))
    views = chord_views(tf)
    outs, st = ctx.render(m, views, mode=V.MODE_FTB, schedule=schedule, term_alpha=1.0)
    assert st["anomalies"] == 0
    for v, g in zip(views, outs):
        check_chord(g, v, n)
    m.close()


def test_gpu_steps_brute_force(ctx):
    print("steps vs brute force", flush=True)
    for (n, seed, variant), v in step_cases():
        inp = _grid(n, seed, variant)
        m = ctx.mesh(inp["x"], inp["t"])
        tf = ctx.tf(inp["tf"])
        v.tf_id = tf
        for sched in (V.SCHED_BATCHED, V.SCHED_PER_VIEW):
            g = ctx.render(m, [v], mode=V.MODE_FTB, schedule=sched, term_alpha=1.0)[0][0]
            check_steps(g["steps"], v, inp)
        m.close()


def test_gpu_ref_steps_brute_force(ctx):
    """REF mode (the reference's walk) on the device against the tet count."""
    print("REF steps vs brute force", flush=True)
    from closed_forms import camera_rays, tets_crossed
    for n in (1, 2, 3, 4):
        inp = _grid(n, 0x5EED0500 + n)
        m = ctx.mesh(inp["x"], inp["t"])
        tf = ctx.tf(inp["tf"])
        v = V.ortho_view(64, 64, tf)
        g = ctx.render(m, [v], mode=V.MODE_REF)[0][0]
        O_, d = camera_rays(v)
        cnt, amb = tets_crossed(O_, d, inp["xn"][:, :3], inp["t"])
        cov = g["entry_ext"] != V.NO_ID
        assert ((cnt > 0) & ~cov & ~amb).sum() == 0
        ok = ~amb & cov
        assert ok.sum() > 0.8 * ok.size
        assert np.array_equal(g["steps"][ok], cnt[ok])
        m.close()


def test_gpu_nonsquare_perspective_equals_oracle(ctx):
    """A non-square perspective view of a jittered grid: entry ids and step
    counts bit for bit, RGBA within 2e-3 of the oracle, PSNR >= 50 dB vs its
    FP64 re-integration."""
    from helpers import oracle_ftb, psnr
    inp = _grid(20, 0x5EED0530)
    m = ctx.mesh(inp["x"], inp["t"])
    tf = ctx.tf(inp["tf"])
    psi = V.psi_generate()
    views = [V.perspective_view(320, 200, tf, eye=(2.0, -1.2, 0.9), fov_y_deg=50.0),
             V.perspective_view(150, 260, tf, eye=(-1.1, 1.9, -1.3), fov_y_deg=70.0)]
    print("non-square perspective 20^3", flush=True)
    outs, _ = ctx.render(m, views, mode=V.MODE_FTB)
    for v, g in zip(views, outs):
        orc = oracle_ftb(inp, v, psi, fp64=True)
        assert orc["anomalies"] == 0
        assert np.array_equal(g["entry_ext"], orc["entry_ext"])
        assert np.array_equal(g["steps"], orc["steps"])
        assert np.abs(g["rgba"] - orc["rgba"]).max() <= 2e-3
        cov = orc["entry_ext"] != V.NO_ID
        assert cov.mean() > 0.2
        assert psnr(g["rgba"], orc["rgba64"], cov) >= 50.0
    m.close()
