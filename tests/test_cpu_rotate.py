"""The CUDA viewer's re-render step (vfRay.cu:617-640): rotateVolume
(vfRayPreComp.cc:886-906) then normalizeVertices, with the limits of the
unrotated mesh kept.  The oracle's restatement is pinned against the
compiled reference's vertices, owner faces and image after two presses of
the space key (tests/golden/ref_rot.npz)."""
from __future__ import annotations

import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import vfray_amd as V
from vfray_amd import workloads as WL
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_rot.npz")


@pytest.fixture(scope="module")
def rot():
    return dict(np.load(GOLD))


def test_gl_rotation_is_rx_times_ry_column_major(rot):
    """vfr_gl_rotation(rx, ry) = Rx(rx) Ry(ry) (glRotatef about x, then about
    y, on the identity), stored column-major; the fixture's matrices are
    those of the same angles."""
    for rx, ry in [(0, 0), (30, 20), (-75, 190), (10, 0), (0, 10)]:
        mv = V.gl_rotation(rx, ry)
        M = mv.reshape(4, 4).T  # column-major -> row-major
        want = Rotation.from_euler("XY", [rx, ry], degrees=True).as_matrix()
        assert np.abs(M[:3, :3] - want).max() < 1e-7
        assert np.array_equal(M[3], [0, 0, 0, 1]) and np.array_equal(M[:3, 3], [0, 0, 0])
    for k, (rx, ry) in enumerate(rot["angles"]):
        assert np.array_equal(V.gl_rotation(float(rx), float(ry)), rot["mv"][k])


def test_rotate_volume_applies_the_transpose():
    """rotateVolume reads the matrix row-major: a column-major GL matrix acts
    as its transpose.  The library's host function equals the oracle bit for
    bit and the rotation in double to rounding."""
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (500, 4)).astype(np.float32)
    mv = V.gl_rotation(33.0, -47.0)
    a, b = V.rotate_volume(x, mv), O.rotate(x, mv)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    M = mv.reshape(4, 4).T.astype(np.float64)
    want = x[:, :3].astype(np.float64) @ M[:3, :3]  # = (M^T v) per row
    assert np.abs(a[:, :3] - want).max() < 1e-6
    assert np.array_equal(a[:, 3], x[:, 3])
    y = x.copy()
    V.lib().vfr_normalize_vertices(V.fptr(y), y.shape[0])
    assert np.array_equal(y.view(np.uint32), O.normalize(x).view(np.uint32))


def test_oracle_rotated_render_matches_reference(rot):
    """Two presses of the space key at (30, 20) on the 8^3 mesh: vertices bit
    for bit, owner faces bit for bit, image within 1 where the reference's
    psi table departs from the closed form, all against the compiled
    reference; the limits are those of the unrotated mesh."""
    x, t, tf, _ = WL.plain_cpu(8)
    xn = O.normalize(x)
    lm = O.lmt(xn, t)
    v = xn
    for mv in rot["mv"]:
        v = O.normalize(O.rotate(v, mv))
    assert np.array_equal(v.view(np.uint32), rot["verts"].view(np.uint32))
    con, _ = O.connectivity(t, x.shape[0])
    r = O.render_ref(v, t, con, O.extf(con), tf, V.psi_generate(), lm, 128, 128)
    assert r["n_vis"] == int(rot["n_vis"][0])
    assert np.array_equal(r["entry_ext"], rot["owner128"])
    d = np.abs(r["rgb8"].astype(int) - rot["rgb128_strict"])
    assert d.max() <= 1 and (d > 0).any(-1).mean() < 0.01
    # the limits went stale: the rotated mesh's own longest edge differs
    assert O.lmt(v, t)[0] != lm[0]
