"""Closed forms and brute force that pin the front-to-back walk without
restating its arithmetic.

Everything here is double precision and built from the view's own parameters
(eye, target, up, fov, image size, or the orthographic matrix), not from the
renderer's view matrix:

  * ``camera_rays``: the world ray through each pixel.  Conventions fixed
    here: pixel (px, py) sits at x = 2 px / (W - 1) - 1, y = 2 py / (H - 1) - 1
    (the reference's pixel-to-object map, vfRayKernelCPU.cc:250, 254); image
    +x is right = forward x up, image +y is up; fov_y spans the image height
    and pixels are square, so the width spans tan(fov_y / 2) W / H.
  * ``slab_chord``: length of a line inside an axis-aligned box.
  * ``line_tet_lengths``: length of a line inside every tet (clipping the
    line against the tet's four face half-spaces), so counting the tets with
    positive length gives the number of tets a ray crosses.
  * ``cube_grid``: an unjittered n^3 hex grid on [0,1]^3, 6 Freudenthal tets
    per cell, so the normalized mesh is exactly the box [-1,1]^3 and its
    longest edge is the cell diagonal 2 sqrt(3) / n.

With a constant transfer function (C, a) the partially pre-integrated
segments telescope: every segment's psi terms cancel (C_f = C_b), each
segment multiplies the transparency by exp(-a dl), and front to back the
colour is C (1 - T) with 1 - T = 1 - exp(-a L / maxEdge), L the ray's chord.
"""
from __future__ import annotations

import numpy as np


def cube_grid(n: int) -> tuple[np.ndarray, np.ndarray]:
    """(xyzs float32 [nV, 4], tets uint32 [6 n^3, 4]); scalar = 0.2 + 0.6 x."""
    g = np.arange(n + 1, dtype=np.float64) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    x = np.concatenate([xyz, (0.2 + 0.6 * xyz[:, :1])], 1).astype(np.float32)

    def vid(i, j, k):
        return (i * (n + 1) + j) * (n + 1) + k

    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for p in perms:
                    c = [i, j, k]
                    ids = [vid(*c)]
                    for ax in p:
                        c = list(c)
                        c[ax] += 1
                        ids.append(vid(*c))
                    tets.append(ids)
    return x, np.array(tets, np.uint32)


def pixel_coords(W: int, H: int) -> tuple[np.ndarray, np.ndarray]:
    return 2.0 * np.arange(W) / (W - 1) - 1.0, 2.0 * np.arange(H) / (H - 1) - 1.0


def camera_rays(view) -> tuple[np.ndarray, np.ndarray]:
    """(origins [H, W, 3], directions [H, W, 3]) of a vfray_amd.View in world
    (normalized mesh) space."""
    W, H = view.width, view.height
    px, py = pixel_coords(W, H)
    if view.projection == 1:
        E = np.array(view.eye[:], np.float64)
        f = np.array(view.target[:], np.float64) - E
        f /= np.linalg.norm(f)
        r = np.cross(f, np.array(view.up[:], np.float64))
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        ty = np.tan(np.radians(float(view.fov_y_deg)) / 2.0)
        tx = ty * W / H
        d = f + (px[None, :, None] * tx) * r + (py[:, None, None] * ty) * u
        return np.broadcast_to(E, d.shape).copy(), d
    # orthographic: (X, Y, Z) = A p + b; pixel (x, y) is the line A^-1 ((x, y, Z) - b)
    M = np.array(view.ortho[:], np.float64).reshape(3, 4)
    Ai = np.linalg.inv(M[:, :3])
    q = np.zeros((H, W, 3))
    q[..., 0] = px[None, :]
    q[..., 1] = py[:, None]
    q[..., 2] = 10.0
    O = (q - M[:, 3]) @ Ai.T
    d = np.broadcast_to(-Ai[:, 2], O.shape).copy()  # towards decreasing Z
    return O, d


def slab_chord(O: np.ndarray, d: np.ndarray, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """Length of the half-line O + s d (s >= 0) inside the box [lo, hi]^3."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (lo - O) / d
        t2 = (hi - O) / d
    tn = np.nanmax(np.minimum(t1, t2), -1)
    tf = np.nanmin(np.maximum(t1, t2), -1)
    return np.maximum(tf - np.maximum(tn, 0.0), 0.0) * np.linalg.norm(d, axis=-1)


def line_tet_lengths(O: np.ndarray, d: np.ndarray, P: np.ndarray, T: np.ndarray) -> np.ndarray:
    """[n_rays, n_tets] length of each half-line O + s d (s >= 0) inside each
    tet (vertices P [nV, 3], tets T [nT, 4]); negative when it misses."""
    O = O.reshape(-1, 3)
    d = d.reshape(-1, 3)
    lo = np.zeros((O.shape[0], T.shape[0]))
    hi = np.full_like(lo, np.inf)
    for f in range(4):
        a, b, c, o = P[T[:, f]], P[T[:, (f + 1) % 4]], P[T[:, (f + 2) % 4]], P[T[:, (f + 3) % 4]]
        nrm = np.cross(b - a, c - a)
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        nrm *= np.sign(((o - a) * nrm).sum(1))[:, None]  # inward
        num = ((O[:, None, :] - a[None]) * nrm[None]).sum(-1)
        den = (d[:, None, :] * nrm[None]).sum(-1)
        with np.errstate(divide="ignore", invalid="ignore"):
            s = -num / den
        lo = np.where(den > 0, np.maximum(lo, s), lo)
        hi = np.where(den < 0, np.minimum(hi, s), hi)
        hi = np.where((den == 0) & (num < 0), -np.inf, hi)
    return (hi - lo) * np.linalg.norm(d, axis=1)[:, None]


def tets_crossed(O, d, P, T, delta: float = 1e-5) -> tuple[np.ndarray, np.ndarray]:
    """(count of tets each ray crosses over a length > delta, ambiguous: some
    tet is met over a length within delta of zero, i.e. the ray grazes an
    edge or a vertex)."""
    shape = O.shape[:-1]
    L = line_tet_lengths(O, d, P.astype(np.float64), T.astype(np.int64))
    return (L > delta).sum(1).reshape(shape), (np.abs(L) < delta).any(1).reshape(shape)


def constant_tf(C, a) -> np.ndarray:
    return np.tile(np.array([C[0], C[1], C[2], a], np.float32), 256)
