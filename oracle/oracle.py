"""ctypes wrapper of the CPU oracle (liboracle.so) and the compiled reference
CPU path (oracle/_ref).  TEST INFRASTRUCTURE: imported only by the tests, the
smoke check and the benchmark's CPU-baseline leg, only to check or time the
renderer against it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle oracle`")
        L = C.CDLL(LIB)
        f32p, u32p, f64p, u8p = (C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_double), C.POINTER(C.c_uint8))
        u32 = C.c_uint32
        L.orc_normalize.argtypes = [f32p, u32]
        L.orc_rotate.argtypes = [f32p, u32, f32p]
        L.orc_connectivity.argtypes = [u32p, u32, u32, u32p]
        L.orc_connectivity.restype = C.c_long
        L.orc_extf.argtypes = [u32p, u32, u32p]
        L.orc_extf.restype = u32
        L.orc_lmt.argtypes = [f32p, u32, u32p, u32, f32p]
        L.orc_psi_exact.argtypes = [C.c_double, C.c_double]
        L.orc_psi_exact.restype = C.c_double
        L.orc_expf.argtypes = [C.c_float]
        L.orc_expf.restype = C.c_float
        L.orc_rcp.argtypes = [C.c_float]
        L.orc_exp_seg.argtypes = [C.c_float]
        L.orc_exp_seg.restype = C.c_float
        L.orc_rcp.restype = C.c_float
        L.orc_psi_series.argtypes = [C.c_float, C.c_float]
        L.orc_psi_series.restype = C.c_float
        L.orc_psi_zeta.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_render_ref.argtypes = [f32p, u32, u32p, u32, u32p, u32p, u32, f32p, f32p, f32p, u32, u32,
                                     u8p, f32p, f64p, u32p, u32p, u32p, u32]
        L.orc_render_ref.restype = C.c_long
        L.orc_render_ftb.argtypes = [f32p, u32, u32p, u32, u32p, u32p, u32, f32p, f32p, f32p, u32, f32p,
                                     C.c_float, C.c_float, u32, u32, C.c_float, f32p, f64p, u32p, u32p,
                                     u32p, u32]
        L.orc_render_ftb.restype = C.c_long
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's render loops; returns the previous
    maximum (n <= 0 only reads it)."""
    return lib().orc_set_threads(n)


def _p(a, t):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(t))


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def u32a(a):
    return np.ascontiguousarray(a, np.uint32)


def normalize(xyzs: np.ndarray) -> np.ndarray:
    x = f32(xyzs).copy()
    lib().orc_normalize(_p(x, C.c_float), x.shape[0])
    return x


def rotate(xyzs: np.ndarray, mv) -> np.ndarray:
    x = f32(xyzs).copy()
    m = f32(mv).reshape(16)
    lib().orc_rotate(_p(x, C.c_float), x.shape[0], _p(m, C.c_float))
    return x


def connectivity(tets: np.ndarray, n_verts: int) -> tuple[np.ndarray, int]:
    t = u32a(tets)
    con = np.zeros((t.shape[0], 4), np.uint32)
    n = lib().orc_connectivity(_p(t, C.c_uint32), t.shape[0], n_verts, _p(con, C.c_uint32))
    if n < 0:
        raise ValueError("a vertex has no incident tetrahedron")
    return con, int(n)


def extf(con: np.ndarray) -> np.ndarray:
    c = u32a(con)
    n = lib().orc_extf(_p(c, C.c_uint32), c.shape[0], None)
    out = np.zeros((max(n, 1), 2), np.uint32)
    lib().orc_extf(_p(c, C.c_uint32), c.shape[0], _p(out, C.c_uint32))
    return out[:n]


def lmt(xyzs: np.ndarray, tets: np.ndarray) -> np.ndarray:
    out = np.zeros(3, np.float32)
    x, t = f32(xyzs), u32a(tets)
    lib().orc_lmt(_p(x, C.c_float), x.shape[0], _p(t, C.c_uint32), t.shape[0], _p(out, C.c_float))
    return out


def expf(x: float) -> float:
    return lib().orc_expf(x)


def exp_seg(x: float) -> float:
    return lib().orc_exp_seg(x)


def rcp(x: float) -> float:
    return lib().orc_rcp(x)


def psi_series(a: float, b: float) -> float:
    return lib().orc_psi_series(a, b)


def psi_zeta(a: float, b: float) -> tuple[float, float]:
    """(psi, zeta) of one segment as the front-to-back walk evaluates them."""
    ps, z = C.c_float(), C.c_float()
    lib().orc_psi_zeta(a, b, C.byref(ps), C.byref(z))
    return ps.value, z.value


def psi_exact(tf: float, tb: float) -> float:
    return lib().orc_psi_exact(tf, tb)


def render_ref(xyzs, tets, con, ext, tf, psi, lmt_, W, H, fp64=False, composite=0) -> dict:
    """main.cc's render (K1, compaction, K2, K3 per face in order) at W x H.
    xyzs must already be normalized.  composite: 0 PPI, 1 linear, 2 linear
    with opacity (fp64 applies to PPI only)."""
    x, t, c, e = f32(xyzs), u32a(tets), u32a(con), u32a(ext)
    tfa, ps, lm = f32(tf).reshape(-1), f32(psi).reshape(-1), f32(lmt_)
    rgb8 = np.zeros((H, W, 3), np.uint8)
    rgba = np.zeros((H, W, 4), np.float32)
    rgba64 = np.zeros((H, W, 4), np.float64) if fp64 else None
    ent = np.zeros((H, W), np.uint32)
    steps = np.zeros((H, W), np.uint32)
    nvis = np.zeros(1, np.uint32)
    r = lib().orc_render_ref(_p(x, C.c_float), x.shape[0], _p(t, C.c_uint32), t.shape[0], _p(c, C.c_uint32),
                             _p(e, C.c_uint32), e.shape[0], _p(tfa, C.c_float), _p(ps, C.c_float),
                             _p(lm, C.c_float), W, H, _p(rgb8, C.c_uint8), _p(rgba, C.c_float),
                             _p(rgba64, C.c_double), _p(ent, C.c_uint32), _p(steps, C.c_uint32),
                             _p(nvis, C.c_uint32), composite)
    if r < 0:
        raise ValueError("bad input")
    return {"rgb8": rgb8, "rgba": rgba, "rgba64": rgba64, "entry_ext": ent, "steps": steps,
            "face_returns": int(r), "n_vis": int(nvis[0])}


def render_ftb(xyzs, tets, con, ext, tf, psi, lmt_, proj, M, k, W, H, term=0.99, fp64=False,
               psi_table=False, composite=0, kx=None) -> dict:
    """The renderer's front-to-back mode restated on the CPU (OpenMP); psi
    from the series, or with psi_table from the table `psi`; composite 1/2:
    the reference's linear integration without / with opacity (rgba = image
    byte / 255 and opacity; rgba64 the same integral in FP64 without the
    per-step truncation).  Perspective: k = tan(fov_y/2), kx = k W/H (None:
    kx = k, a square view)."""
    x, t, c, e = f32(xyzs), u32a(tets), u32a(con), u32a(ext)
    tfa, lm, Mm = f32(tf).reshape(-1), f32(lmt_), f32(M).reshape(-1)
    ps = f32(psi).reshape(-1) if psi_table else None
    rgba = np.zeros((H, W, 4), np.float32)
    rgba64 = np.zeros((H, W, 4), np.float64) if fp64 else None
    ent = np.zeros((H, W), np.uint32)
    steps = np.zeros((H, W), np.uint32)
    nent = np.zeros((H, W), np.uint32)
    r = lib().orc_render_ftb(_p(x, C.c_float), x.shape[0], _p(t, C.c_uint32), t.shape[0], _p(c, C.c_uint32),
                             _p(e, C.c_uint32), e.shape[0], _p(tfa, C.c_float), _p(ps, C.c_float),
                             _p(lm, C.c_float), proj, _p(Mm, C.c_float), k, k if kx is None else kx, W, H,
                             term,
                             _p(rgba, C.c_float), _p(rgba64, C.c_double), _p(ent, C.c_uint32),
                             _p(steps, C.c_uint32), _p(nent, C.c_uint32), composite)
    if r < 0:
        raise ValueError("bad input")
    return {"rgba": rgba, "rgba64": rgba64, "entry_ext": ent, "steps": steps, "n_entries": nent,
            "anomalies": int(r)}


# ------------------------------------------------------------ compiled reference

def ref_available(flavour: str = "strict") -> bool:
    return os.path.exists(os.path.join(REF_DIR, flavour, "vfRayCPU"))


def run_ref_cli(flavour: str, work: str, name: str) -> str:
    """Runs the reference's own vfRayCPU (main.cc) on <work>/tet_offs/<name>.off
    from <work>/run; returns the run directory (PPM there, caches in tet_offs)."""
    run = os.path.join(work, "run")
    os.makedirs(run, exist_ok=True)
    exe = os.path.join(REF_DIR, flavour, "vfRayCPU")
    r = subprocess.run([exe, name], cwd=run, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference vfRayCPU failed:\n{r.stdout}")
    return run


def run_ref_harness(flavour: str, prefix: str, W: int, H: int, rotations=None) -> dict:
    """Per-pixel last-writer face ids and the image from the reference CPU
    path (see ref_harness.cpp); rotations: k 16-float matrices applied as the
    viewer's rotate + renormalize before rendering."""
    exe = os.path.join(REF_DIR, flavour, "ref_harness")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "o")
        args = [exe, prefix, str(W), str(H), out]
        if rotations is not None:
            rf = os.path.join(td, "rot.bin")
            np.ascontiguousarray(rotations, np.float32).tofile(rf)
            args.append(rf)
        r = subprocess.run(args, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness failed:\n{r.stdout}")
        rgb = np.fromfile(out + ".rgb", np.uint8).reshape(H, W, 3)
        owner = np.fromfile(out + ".owner", np.uint32).reshape(H, W)
        info = dict(l.split() for l in open(out + ".txt").read().splitlines())
        verts = np.fromfile(out + ".verts", np.float32).reshape(-1, 4)
    return {"rgb8": rgb, "entry_ext": owner, "verts": verts,
            "info": {k: int(v) for k, v in info.items()}}
