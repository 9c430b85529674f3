"""Shared inputs and comparisons for the parity tests."""
from __future__ import annotations

import numpy as np

import vfray_amd as V
from vfray_amd import workloads as WL
import oracle as O


def c1_inputs(n: int = 16):
    """The 512^2 orthographic workload: rotated mesh, its normalized copy,
    oracle connectivity and limits, TF."""
    x, t, tf, side = WL.plain_cpu(n)
    xn = O.normalize(x)
    con, _ = O.connectivity(t, x.shape[0])
    ext = O.extf(con)
    lm = O.lmt(xn, t)
    return dict(x=x, t=t, tf=tf, side=side, xn=xn, con=con, ext=ext, lmt=lm)


def grid_inputs(n: int, seed: int, variant: int = 0):
    x, t = V.synth_grid(n, seed, variant)
    xn = O.normalize(x)
    con, _ = O.connectivity(t, x.shape[0])
    ext = O.extf(con)
    lm = O.lmt(xn, t)
    return dict(x=x, t=t, tf=V.synth_tf(seed), xn=xn, con=con, ext=ext, lmt=lm)


def psnr(a: np.ndarray, b: np.ndarray, mask: np.ndarray) -> float:
    """10 log10(1/MSE) over RGBA in [0,1] on the masked pixels."""
    d = (a[mask].astype(np.float64) - b[mask].astype(np.float64))
    mse = float((d * d).mean()) if d.size else 0.0
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)


def oracle_ftb(inp, view: V.View, psi, term=0.99, fp64=False, psi_table=False, composite=0):
    proj, M, k, kx = V.view_matrix(view)
    return O.render_ftb(inp["xn"], inp["t"], inp["con"], inp["ext"], inp["tf"], psi, inp["lmt"],
                        proj, M, k, view.width, view.height, term=term, fp64=fp64, psi_table=psi_table,
                        composite=composite, kx=kx)


def legacy_tf_text(seed: int = 7) -> str:
    """A legacy transfer-function file as cleanTF reads it: 7 header lines,
    then 256 lines "index r g b a" with 17-digit values (seeded)."""
    rng = np.random.default_rng(seed)
    head = ["# legacy transfer function", "# produced for the tests", "format 1", "entries 256",
            "columns index r g b a", "range 0 1", "#"]
    vals = rng.random((256, 4))
    vals[rng.random((256, 4)) < 0.1] = 0.0
    lines = [f"{i} " + " ".join(repr(float(v)) for v in vals[i]) for i in range(256)]
    return "\n".join(head + lines) + "\n"


def brute_con(t: np.ndarray) -> np.ndarray:
    """Neighbour = smallest other tet holding all three face vertices (O(T^2)),
    the reference's rule (vfRayPreComp.cc:219-270) by exhaustive search."""
    nT = t.shape[0]
    sets = [set(map(int, r)) for r in t]
    con = np.zeros((nT, 4), np.uint32)
    for i in range(nT):
        for f in range(4):
            fv = {int(t[i, (k + f) % 4]) for k in range(3)}
            cand = [j for j in range(nT) if j != i and fv <= sets[j]]
            con[i, f] = min(cand) if cand else i
    return con


CONNECTIVITY_CASES = ("one_tet", "cube", "grid2", "nonmanifold", "grid3_jitter")


def connectivity_case(case: str):
    """Vertices (nV x 4: x, y, z, s) and tets of a small connectivity case;
    "nonmanifold" puts three tets on the face (0, 1, 2)."""
    if case == "one_tet":
        t = np.array([[0, 1, 2, 3]], np.uint32)
        x = np.array([[0, 0, 0, 0.1], [1, 0, 0, 0.2], [0, 1, 0, 0.3], [0, 0, 1, 0.4]], np.float32)
    elif case == "nonmanifold":
        t = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [2, 1, 0, 5], [3, 4, 1, 2]], np.uint32)
        x = np.array([[0, 0, 0, 0.1], [1, 0, 0, 0.2], [0, 1, 0, 0.3], [0.2, 0.2, 1, 0.4],
                      [0.3, 0.3, -1, 0.5], [0.25, 0.2, 0.5, 0.6]], np.float32)
    else:
        n = {"cube": 1, "grid2": 2, "grid3_jitter": 3}[case]
        x, t = V.synth_grid(n, 11, 0)
    return x, t

