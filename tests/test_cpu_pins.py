"""Pins of the oracle's front-to-back walk that do not restate its formulas.

* Colour: on an unjittered cube with a constant transfer function, every
  covered pixel's opacity is 1 - exp(-a L / maxEdge) and its colour C times
  that, L the pixel ray's chord through the box by a double-precision slab
  test, the ray built from eye/target/up/fov and the image size alone
  (closed_forms.camera_rays).  Square and non-square perspective views and an
  orthographic one with a scaled, rotated matrix.
* Step counts: the number of tets the walk visits equals the number of tets
  the pixel's line crosses by double line-tet clipping, on jittered 1^3-4^3
  grids (front-to-back perspective and the reference's own orthographic walk)
  and on voided meshes whose rays exit and re-enter.
* The chord pin has teeth: planted mistakes in the oracle's camera inputs (a
  fov_y-only aspect, the full fov in place of its half, the per-pixel ray
  length factor dropped) fail it.
"""
import numpy as np
import pytest

import vfray_amd as V
from vfray_amd import workloads as WL
import oracle as O
from closed_forms import camera_rays, constant_tf, cube_grid, slab_chord, tets_crossed

C_RGB = np.array([0.8, 0.5, 0.3], np.float32)
ALPHA = np.float32(0.15)
CHORD_TOL = 1e-5   # observed below 8e-7 (FP32 walk vs double closed form)


@pytest.fixture(scope="module")
def cube8():
    n = 8
    x, t = cube_grid(n)
    xn = O.normalize(x)
    con, _ = O.connectivity(t, x.shape[0])
    return dict(n=n, x=x, t=t, xn=xn, con=con, ext=O.extf(con), lmt=O.lmt(xn, t),
                tf=constant_tf(C_RGB, ALPHA))


def chord_views(tf_id: int = 0):
    eye = np.array([1.7, -1.2, 1.3])
    eye = 2.5 * eye / np.linalg.norm(eye)
    eye2 = np.array([-0.4, 2.1, 1.2])
    eye2 = 2.6 * eye2 / np.linalg.norm(eye2)
    c, s = np.cos(0.5), np.sin(0.5)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    A = 0.55 * R
    ortho = np.hstack([A, np.array([[0.02], [-0.03], [0.0]])]).astype(np.float32)
    return [V.perspective_view(64, 64, tf_id, eye=eye, fov_y_deg=60.0),
            V.perspective_view(96, 64, tf_id, eye=eye, fov_y_deg=60.0),
            V.perspective_view(48, 80, tf_id, eye=eye2, fov_y_deg=45.0),
            V.ortho_view(72, 56, tf_id, matrix=ortho)]


def check_chord(g: dict, view, n: int) -> None:
    """g: a rendered view (rgba, entry_ext) of the normalized cube8 with the
    constant TF and no early termination."""
    O_, d = camera_rays(view)
    L = slab_chord(O_, d)
    op = 1.0 - np.exp(-float(ALPHA) * L / (2.0 * np.sqrt(3.0) / n))
    cov = g["entry_ext"] != V.NO_ID
    inside = L > 1e-2
    # coverage is the box's silhouette, up to rays grazing its edges
    assert (inside & ~cov).sum() == 0
    assert (cov & ~(L > 0)).sum() <= 0.002 * cov.size
    sel = inside & cov
    assert sel.sum() > 0.3 * sel.size
    assert np.abs(g["rgba"][..., 3] - op)[sel].max() <= CHORD_TOL
    assert np.abs(g["rgba"][..., :3] - C_RGB * op[..., None])[sel].max() <= CHORD_TOL


def oracle_view(inp, view, psi, term=1.0):
    proj, M, k, kx = V.view_matrix(view)
    return O.render_ftb(inp["xn"], inp["t"], inp["con"], inp["ext"], inp["tf"], psi, inp["lmt"], proj, M, k,
                        view.width, view.height, term=term, kx=kx)


def test_chord_closed_form(cube8):
    psi = V.psi_generate()
    for v in chord_views():
        check_chord(oracle_view(cube8, v, psi), v, cube8["n"])


@pytest.mark.parametrize("mistake", ["fov_y_both_axes", "full_fov", "ray_factor_dropped"])
def test_chord_pin_rejects_planted_mistakes(cube8, mistake):
    """The same check on oracle renders given a wrong camera: each fails."""
    psi = V.psi_generate()
    v = chord_views()[1]  # 96 x 64
    proj, M, k, kx = V.view_matrix(v)
    M = M.copy()
    if mistake == "fov_y_both_axes":
        M[0:4] *= np.float32(kx / k)
        kx = k
    elif mistake == "full_fov":
        s = np.tan(np.radians(30.0)) / np.tan(np.radians(60.0))
        M[0:8] *= np.float32(s)
        k, kx = k / s, kx / s
    else:
        k = kx = 0.0
    g = O.render_ftb(cube8["xn"], cube8["t"], cube8["con"], cube8["ext"], cube8["tf"], psi, cube8["lmt"], proj, M,
                     k, v.width, v.height, term=1.0, kx=kx)
    with pytest.raises(AssertionError):
        check_chord(g, v, cube8["n"])


def _grid(n, seed, variant=0):
    x, t = V.synth_grid(n, seed, variant)
    xn = O.normalize(x)
    con, _ = O.connectivity(t, x.shape[0])
    return dict(x=x, t=t, xn=xn, con=con, ext=O.extf(con), lmt=O.lmt(xn, t), tf=V.synth_tf(seed))


def step_cases():
    """(mesh, view) pairs of the step-count pin: jittered 1^3-4^3 grids in
    non-square perspective views, and voided 6^3 / 8^3 meshes."""
    out = []
    for n in (1, 2, 3, 4):
        v = WL.sphere_views(0x5EED0510 + n, 1, [64], 0)[0]
        v.width, v.height = 80, 56
        out.append(((n, 0x5EED0500 + n, 0), v))
    for n in (6, 8):
        v = WL.sphere_views(0x5EED0520 + n, 1, [64], 0)[0]
        v.width, v.height = 72, 60
        out.append(((n, 0x5EED0303, 1), v))
    return out


def check_steps(steps: np.ndarray, view, inp, min_checked: float = 0.9) -> None:
    O_, d = camera_rays(view)
    cnt, amb = tets_crossed(O_, d, inp["xn"][:, :3], inp["t"])
    ok = ~amb
    assert ok.mean() >= min_checked
    assert (cnt[ok] > 0).sum() > 0.3 * ok.sum()
    assert np.array_equal(steps[ok], cnt[ok])


@pytest.mark.parametrize("case", range(6))
def test_ftb_steps_brute_force(case):
    (n, seed, variant), v = step_cases()[case]
    inp = _grid(n, seed, variant)
    psi = V.psi_generate()
    r = oracle_view(inp, v, psi, term=1.0)
    if variant == 1:
        assert (r["n_entries"] > 2).sum() > 10  # rays that exit and re-enter
    check_steps(r["steps"], v, inp)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_ref_steps_brute_force(n):
    """The reference's orthographic walk along -z (owner faces, its exit
    rule with the RDELTA slack): steps = tets crossed on every pixel it
    renders whose line grazes no edge or vertex."""
    inp = _grid(n, 0x5EED0500 + n)
    psi = V.psi_generate()
    r = O.render_ref(inp["xn"], inp["t"], inp["con"], inp["ext"], inp["tf"], psi, inp["lmt"], 64, 64)
    v = V.ortho_view(64, 64, 0)
    O_, d = camera_rays(v)
    cnt, amb = tets_crossed(O_, d, inp["xn"][:, :3], inp["t"])
    cov = r["entry_ext"] != V.NO_ID
    assert ((cnt > 0) & ~cov & ~amb).sum() == 0
    ok = ~amb & cov
    assert ok.sum() > 0.8 * ok.size
    assert np.array_equal(r["steps"][ok], cnt[ok])
