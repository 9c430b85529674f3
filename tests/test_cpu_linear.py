"""Linear compositing (the reference's USEPARTIALPREINT 0 build,
vfRayKernel.cu:738-806) in the CPU oracle, pinned by the integrals it
approximates: per segment, colour interpolated linearly between the two TF
samples times (1 - opacity accumulated linearly along the segment), so one
tet reproduces the quadrature of that integrand, and over a ray with a
constant TF the per-segment sums telescope to 255 c (L - a L^2 / 2) for a
thickness L."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.integrate import quad

import vfray_amd as V
import oracle as O
from helpers import grid_inputs


@pytest.fixture(scope="module")
def psi():
    return V.psi_generate()


def _one_tet():
    x = np.array([[0, 0, 1, 0], [1, 0, 1, 0], [0, 1, 1, 0], [0.2, 0.2, 0, 0]], np.float32)
    t = np.array([[0, 1, 2, 3]], np.uint32)
    xn = O.normalize(x)
    xn[:, 3] = np.array([0.12, 0.55, 0.83, 0.31], np.float32)
    con, _ = O.connectivity(t, 4)
    return xn, t, con, O.extf(con), O.lmt(xn, t)


def _ramp_tf():
    i = np.arange(256) / 255.0
    tf = np.stack([0.2 + 0.7 * i, 0.9 - 0.8 * i, 0.5 + 0.4 * np.sin(6 * i), 0.05 + 0.6 * i], 1)
    return tf.astype(np.float32).reshape(-1)


@pytest.mark.parametrize("composite", [1, 2])
def test_linear_single_tet_matches_quadrature(psi, composite):
    """One tet, scalar linear in space, ramp TF: the colour byte is the
    integral of c(u) (1 - O(u)) over the segment (c linear between the entry
    and exit TF samples, O(u) the integral of the linearly interpolated
    alpha; O = 0 without opacity), truncated; the opacity is the integral of
    alpha; the image byte is colour * opacity."""
    xn, t, con, ext, lm = _one_tet()
    tf = _ramp_tf()
    W = 64
    r = O.render_ref(xn, t, con, ext, tf, psi, lm, W, W, composite=composite)
    ys, xs = np.nonzero((r["entry_ext"] != V.NO_ID) & (r["steps"] >= 1))
    assert ys.size > 1000
    P = xn[:, :3].astype(np.float64)
    S = xn[:, 3].astype(np.float64)
    tfa = tf.reshape(256, 4).astype(np.float64)
    tf32 = tf.reshape(256, 4)
    # barycentric coordinates of a point in the tet, from the 3x3 system
    Minv = np.linalg.inv(np.stack([P[0] - P[3], P[1] - P[3], P[2] - P[3]], 1))

    def s_at(p):
        lam = Minv @ (p - P[3])
        return lam[0] * S[0] + lam[1] * S[1] + lam[2] * S[2] + (1 - lam.sum()) * S[3]

    def zplane(a, b, c):
        n = np.cross(b - a, c - a)
        return lambda X, Y: a[2] - (n[0] * (X - a[0]) + n[1] * (Y - a[1])) / n[2]
    top = zplane(P[0], P[1], P[2])
    lower = [zplane(P[1], P[2], P[3]), zplane(P[2], P[3], P[0]), zplane(P[3], P[0], P[1])]
    checked = 0
    for y, x in zip(ys, xs):
        ox = float(np.float32(x * 2) / np.float32(W - 1) - np.float32(1))
        oy = float(np.float32(y * 2) / np.float32(W - 1) - np.float32(1))
        z0 = top(ox, oy)
        z1 = max(f(ox, oy) for f in lower)
        dz = z0 - z1
        s0, s1 = s_at(np.array([ox, oy, z0])), s_at(np.array([ox, oy, z1]))
        f0, f1 = 255 * s0, 255 * s1
        if min(abs(f0 - round(f0)), abs(f1 - round(f1))) < 1e-3 or dz < 1e-3:
            continue  # TF sample index or step ambiguous at float precision
        i0, i1 = int(f0), int(f1)
        for ch in range(3):
            cP, cN = 255 * float(tf32[i0, ch]), 255 * float(tf32[i1, ch])
            aP, aN = (float(tf32[i0, 3]), float(tf32[i1, 3])) if composite == 2 else (0.0, 0.0)

            def c(u):
                return cP + (cN - cP) * u / dz

            def O_(u):
                return aP * u + (aN - aP) * u * u / (2 * dz)
            integral = quad(lambda u: c(u) * (1 - O_(u)), 0, dz, epsabs=1e-12, epsrel=1e-12)[0]
            col = min(int(np.floor(integral)), 255)
            op = quad(lambda u: aP + (aN - aP) * u / dz, 0, dz)[0] if composite == 2 else 1.0
            frac = integral - np.floor(integral)
            if 1e-4 < frac < 1 - 1e-4:
                opf = min(op, 1.0)
                want = int(np.float32(col) * np.float32(opf)) if composite == 2 else col
                got = int(r["rgb8"][y, x, ch])
                if composite == 2 and abs(col * opf - round(col * opf)) < 1e-3:
                    assert abs(got - want) <= 1
                else:
                    assert got == want, (y, x, ch, got, want, integral, op)
                checked += 1
        if composite == 2:
            assert abs(r["rgba"][y, x, 3] - min(op, 1.0)) < 1e-5
    assert checked > 2000
    assert np.array_equal(tfa.shape, (256, 4))


def test_linear_constant_tf_telescopes(psi):
    """Axis-aligned 8^3 grid (normalized to the [-1,1] cube, every covered
    pixel crosses thickness L = 2), constant TF: without opacity the colour
    is 255 c L less what the per-step truncation drops (under 1 per step);
    with opacity a and a L <= 1 the steps sum to 255 c (L - a L^2 / 2) and the
    opacity to a L; with a L > 1 the ray stops once the opacity passes 1,
    where L - L^2 / 2 peaks, at 1/2."""
    g = grid_inputs(8, 0x5EED0007)
    z = g["xn"][:, 2]
    assert abs(z.max() - 1) < 1e-6 and abs(z.min() + 1) < 1e-6
    L = 2.0
    # rays that cross the whole cube: in PPI mode with constant alpha their
    # FP64 transmittance is exp(-alpha L / maxEdge) (rays through pixels on
    # face edges end early on the reference's exit-test slack)
    tfp = np.tile(np.array([0.5, 0.5, 0.5, 0.3], np.float32), 256)
    rp = O.render_ref(g["xn"], g["t"], g["con"], g["ext"], tfp, psi, g["lmt"], 96, 96, fp64=True)
    whole = np.abs(rp["rgba64"][..., 3] - (1 - np.exp(-0.3 * L / float(g["lmt"][0])))) < 1e-6
    assert whole.sum() > 0.9 * (rp["entry_ext"] != V.NO_ID).sum() > 5000
    cases = [(1, 0.2, 0.0), (2, 0.3, 0.45), (2, 0.35, 1.0)]
    for comp, c, a in cases:
        tf = np.tile(np.array([c, c / 2, c / 4, a], np.float32), 256)
        r = O.render_ref(g["xn"], g["t"], g["con"], g["ext"], tf, psi, g["lmt"], 96, 96, composite=comp)
        cov = whole
        steps = r["steps"][cov].astype(np.float64)
        rgb = r["rgb8"][cov].astype(np.float64)
        alpha = r["rgba"][cov][:, 3].astype(np.float64)
        for ch, cc in enumerate([c, c / 2, c / 4]):
            if comp == 1:
                hi = 255 * np.float32(cc) * L
                assert (rgb[:, ch] <= hi + 1e-3).all() and (rgb[:, ch] >= hi - steps - 1e-3).all()
            elif a * L <= 1:
                hi = 255 * np.float32(cc) * (L - a * L * L / 2)
                col_lo = hi - steps - 1e-3
                # image byte = colour * opacity, truncated
                assert (rgb[:, ch] <= hi * a * L + 1e-3).all()
                assert (rgb[:, ch] >= np.floor(np.maximum(col_lo, 0) * a * L) - 1).all()
            else:
                hi = 255 * np.float32(cc) * 0.5
                assert (rgb[:, ch] <= hi + 1e-3).all() and (rgb[:, ch] >= hi - steps - 1).all()
        if comp == 1:
            assert (alpha == 1.0).all()
        elif a * L <= 1:
            assert np.abs(alpha - a * L).max() < 1e-5
        else:
            assert (alpha == 1.0).all()
            # the early-out stops the walk about half way
            full = O.render_ref(g["xn"], g["t"], g["con"], g["ext"], tf, psi, g["lmt"], 96, 96, composite=1)
            assert (r["steps"][cov] < full["steps"][cov]).mean() > 0.9


def test_linear_entry_and_steps_follow_the_ref_walk(psi):
    """Compositing does not change the walk: entry faces equal PPI mode's and
    so do step counts, except where the opacity passed 1 (composite 2; with
    alpha this small only at pixels on face edges, whose near-degenerate
    planes give a huge first dz)."""
    g = grid_inputs(8, 0x5EED0003)
    tf = g["tf"].copy().reshape(256, 4)
    tf[:, 3] *= 0.05
    tf = tf.reshape(-1)
    r0 = O.render_ref(g["xn"], g["t"], g["con"], g["ext"], tf, psi, g["lmt"], 128, 128)
    for comp in (1, 2):
        r = O.render_ref(g["xn"], g["t"], g["con"], g["ext"], tf, psi, g["lmt"], 128, 128, composite=comp)
        assert np.array_equal(r["entry_ext"], r0["entry_ext"])
        ran = r["rgba"][..., 3] < 1.0 if comp == 2 else np.ones(r["steps"].shape, bool)
        assert ran.mean() > 0.99
        assert np.array_equal(r["steps"][ran], r0["steps"][ran])
        assert r["rgb8"].max() > 0


# ------------------------------------------------ front-to-back (FTB) linear


def _views(seed, n, side):
    from vfray_amd import workloads as WL
    return WL.sphere_views(seed, n, [side], 0)


@pytest.mark.parametrize("composite", [1, 2])
def test_ftb_linear_against_fp64(psi, composite):
    """Linear compositing on the front-to-back walk, perspective views: each
    colour byte is the FP64 integral of the same segments (no per-step
    truncation, clamped at 255) less under 1 per step it composited; the
    opacity is the FP64 opacity.  The walk itself is PPI mode's without
    termination: same entry faces, same steps where the opacity stop did not
    end the ray."""
    from helpers import oracle_ftb
    g = grid_inputs(16, 0x5EED0311)
    tf = g["tf"].copy().reshape(256, 4)
    tf[:, 3] *= 0.6 if composite == 2 else 1.0
    g = dict(g, tf=tf.reshape(-1))
    checked = 0
    for v in _views(0x5EED0311, 2, 160):
        r0 = oracle_ftb(g, v, psi, term=1.0)
        r = oracle_ftb(g, v, psi, fp64=True, composite=composite)
        cov = r["entry_ext"] != V.NO_ID
        assert r["anomalies"] == 0 and cov.mean() > 0.3
        assert np.array_equal(r["entry_ext"], r0["entry_ext"])
        op, op64 = r["rgba"][..., 3].astype(np.float64), r["rgba64"][..., 3]
        ran = (op < 1.0) if composite == 2 else cov
        assert np.array_equal(r["steps"][ran], r0["steps"][ran])
        steps = r["steps"][cov].astype(np.float64)
        b32 = np.rint(r["rgba"][cov][:, :3].astype(np.float64) * 255.0)
        c64 = r["rgba64"][cov][:, :3] * 255.0
        if composite == 1:
            assert (op[cov] == 1.0).all()
            assert (b32 <= c64 + 1e-3).all()
            assert (b32 >= c64 - steps[:, None] - 1e-3).all()
        else:
            assert np.abs(op[cov] - op64[cov]).max() < 2e-4
            o = op64[cov][:, None]
            assert (b32 <= c64 + 1e-2).all()
            assert (b32 >= c64 - (steps[:, None] + 1) * o - 1.01).all()
        assert b32.max() > 50
        checked += int(cov.sum())
    assert checked > 10000


def test_ftb_linear_constant_tf_telescopes(psi):
    """Orthographic FTB view straight down the axis-aligned 8^3 cube (every
    covered pixel crosses thickness L = 2 in object units), constant TF
    without opacity: the colour is 255 c L less what the per-step
    truncation drops (under 1 per step), as on the REF walk."""
    from helpers import oracle_ftb
    g = grid_inputs(8, 0x5EED0007)
    c = 0.2
    g = dict(g, tf=np.tile(np.array([c, c / 2, c / 4, 0.0], np.float32), 256))
    v = V.ortho_view(96, 96, 0)
    r = oracle_ftb(g, v, psi, composite=1)
    cov = r["entry_ext"] != V.NO_ID
    inner = np.zeros_like(cov)
    inner[8:-8, 8:-8] = True
    sel = cov & inner
    assert sel.sum() > 4000
    steps = r["steps"][sel].astype(np.float64)
    rgb = np.rint(r["rgba"][sel][:, :3].astype(np.float64) * 255.0)
    L = 2.0
    for ch, cc in enumerate([c, c / 2, c / 4]):
        hi = 255 * np.float32(cc) * L
        assert (rgb[:, ch] <= hi + 1e-3).all() and (rgb[:, ch] >= hi - steps - 1e-3).all()


def test_ftb_linear_ortho_close_to_ref_linear(psi):
    """The same orthographic 512^2 view on the 16^3 mesh in REF and FTB
    linear modes: the two walks cross the same tets with segment lengths
    computed two ways (plane differences vs weighted face values), so bytes
    agree to within a few units where both own the pixel with the same
    entry face."""
    from helpers import c1_inputs, oracle_ftb
    g = c1_inputs()
    v = V.ortho_view(512, 512, 0)
    for comp in (1, 2):
        rf = oracle_ftb(g, v, psi, composite=comp)
        rr = O.render_ref(g["xn"], g["t"], g["con"], g["ext"], g["tf"], psi, g["lmt"], 512, 512, composite=comp)
        both = (rf["entry_ext"] != V.NO_ID) & (rr["entry_ext"] == rf["entry_ext"])
        assert both.mean() > 0.5
        bf = np.rint(rf["rgba"][both][:, :3].astype(np.float64) * 255.0)
        br = rr["rgb8"][both].astype(np.float64)
        d = np.abs(bf - br)
        assert np.median(d) <= 1 and np.mean(d <= 3) > 0.99, (comp, np.median(d), np.mean(d <= 3))
