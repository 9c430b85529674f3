"""vfray_amd -- Python side of the MI355X-native VF-Ray renderer.

Everything here is a thin ctypes layer over libvfr.so (include/vfr.h); the
renderer itself is C++ host code and gfx950 HIP kernels.  Loading fails loudly
when the library has not been built: there is no Python fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvfr.so")

VFR_OK = 0
PROJ_ORTHO, PROJ_PERSPECTIVE = 0, 1
MODE_REF, MODE_FTB = 0, 1
COMPOSITE_PPI, COMPOSITE_LINEAR, COMPOSITE_LINEAR_OPACITY = 0, 1, 2
RENDER_PERSISTENT = 0x1
RENDER_PIPELINED = 0x2
RENDER_PSI_TABLE = 0x4
RENDER_WIDE = 0x8
RENDER_WAVE_QUEUE = 0x10
RENDER_LDS_RECORDS = 0x20
RENDER_FACE_MEMO = 0x40
SCHED_PER_VIEW, SCHED_BATCHED = 0, 1
SERVE_SLOTS, SERVE_PIPELINE = 0, 1
MESH_NORMALIZE, MESH_MORTON = 0x1, 0x2
NO_ID = 0xFFFFFFFF


class VfrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vfr error {code}: {msg}")
        self.code = code


class View(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("projection", C.c_uint32),
                ("tf_id", C.c_int32), ("ortho", C.c_float * 12), ("eye", C.c_float * 3),
                ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov_y_deg", C.c_float)]


class RenderOpts(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("schedule", C.c_uint32), ("term_alpha", C.c_float),
                ("waves_per_cu", C.c_uint32), ("refill", C.c_uint32), ("use_graph", C.c_uint32),
                ("max_entries", C.c_uint32), ("flags", C.c_uint32), ("occupancy", C.c_uint32),
                ("composite", C.c_uint32)]


class FrameOut(C.Structure):
    _fields_ = [("rgba", C.POINTER(C.c_float)), ("rgb8", C.POINTER(C.c_uint8)),
                ("entry_ext", C.POINTER(C.c_uint32)), ("entry_tet", C.POINTER(C.c_uint32)),
                ("steps", C.POINTER(C.c_uint32))]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("active_rays", C.c_uint64), ("steps", C.c_uint64),
                ("anomalies", C.c_uint64), ("ms_total", C.c_double), ("ms_visibility", C.c_double),
                ("ms_traversal", C.c_double), ("ms_frame_p50", C.c_double),
                ("ms_frame_p99", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerCfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("streams", C.c_uint32), ("max_batch", C.c_uint32),
                ("term_alpha", C.c_float), ("flags", C.c_uint32), ("occupancy", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("refill", C.c_uint32), ("composite", C.c_uint32),
                ("max_batch_pixels", C.c_uint64), ("schedule", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("keep_image", C.c_uint32), ("view", View)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("batch_size", C.c_uint32), ("latency_ms", C.c_double), ("queue_ms", C.c_double)]


class Report(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("rays", C.c_uint64), ("active_rays", C.c_uint64),
                ("steps", C.c_uint64), ("anomalies", C.c_uint64), ("batches", C.c_uint64),
                ("makespan_s", C.c_double), ("mrays_per_s", C.c_double), ("latency_p50_ms", C.c_double),
                ("latency_p90_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_max_ms", C.c_double), ("mean_batch", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshInfo(C.Structure):
    _fields_ = [("n_verts", C.c_uint32), ("n_tets", C.c_uint32), ("n_ext_faces", C.c_uint32),
                ("max_edge_length", C.c_float), ("max_z", C.c_float), ("min_z", C.c_float),
                ("device_bytes", C.c_uint64), ("build_ms", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    """The loaded libvfr.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m vfray_amd.build`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u32, f32p, u32p = C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        sig = {
            "vfr_last_error": (C.c_char_p, []),
            "vfr_version": (C.c_char_p, []),
            "vfr_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
            "vfr_ctx_destroy": (None, [P]),
            "vfr_psi_load": (C.c_int, [P, f32p]),
            "vfr_psi_read": (C.c_int, [P, f32p]),
            "vfr_psi_generate": (C.c_int, [f32p]),
            "vfr_tf_create": (C.c_int, [P, f32p, C.POINTER(C.c_int)]),
            "vfr_bandwidth_probe": (C.c_int, [P, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
            "vfr_mesh_create": (C.c_int, [P, f32p, u32, u32p, u32, u32p, u32p, u32, f32p, u32, C.POINTER(P)]),
            "vfr_mesh_destroy": (None, [P]),
            "vfr_mesh_get_info": (C.c_int, [P, C.POINTER(MeshInfo)]),
            "vfr_mesh_connectivity": (C.c_int, [P, u32p, u32p, f32p]),
            "vfr_mesh_vertices": (C.c_int, [P, f32p]),
            "vfr_mesh_rotate": (C.c_int, [P, f32p, u32]),
            "vfr_plan_create": (C.c_int, [P, P, C.POINTER(View), C.c_int, C.POINTER(RenderOpts), C.POINTER(P)]),
            "vfr_plan_execute": (C.c_int, [P, C.POINTER(Stats)]),
            "vfr_plan_read": (C.c_int, [P, C.c_int, C.POINTER(FrameOut)]),
            "vfr_plan_destroy": (None, [P]),
            "vfr_plan_frame_times": (C.c_int, [P, C.POINTER(C.c_double)]),
            "vfr_plan_autotune": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
            "vfr_render": (C.c_int, [P, P, C.POINTER(View), C.c_int, C.POINTER(RenderOpts), C.POINTER(FrameOut), C.POINTER(Stats)]),
            "vfr_render_noalloc": (C.c_int, [P, P, C.POINTER(View), C.c_int, C.POINTER(RenderOpts),
                                             C.POINTER(FrameOut), C.POINTER(Stats)]),
            "vfr_free": (None, [P]),
            "vfr_server_create": (C.c_int, [P, C.POINTER(ServerCfg), C.POINTER(P), C.c_int, C.POINTER(P)]),
            "vfr_server_submit": (C.c_int, [P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
            "vfr_server_poll": (C.c_int, [P, C.c_uint64, C.POINTER(Result)]),
            "vfr_server_wait": (C.c_int, [P, C.c_uint64, C.POINTER(Result), f32p]),
            "vfr_server_replay": (C.c_int, [P, C.POINTER(Request), C.POINTER(C.c_double), C.c_int,
                                            C.POINTER(Report)]),
            "vfr_server_destroy": (None, [P]),
            "vfr_read_off": (C.c_int, [C.c_char_p, C.POINTER(f32p), u32p, C.POINTER(u32p), u32p]),
            "vfr_write_off": (C.c_int, [C.c_char_p, f32p, u32, u32p, u32]),
            "vfr_normalize_vertices": (None, [f32p, u32]),
            "vfr_rotate_volume": (None, [f32p, u32, f32p]),
            "vfr_gl_rotation": (None, [C.c_float, C.c_float, f32p]),
            "vfr_clean_tf": (C.c_int, [C.c_char_p, C.c_char_p]),
            "vfr_read_con": (C.c_int, [C.c_char_p, u32, C.POINTER(u32p), u32p]),
            "vfr_read_incid": (C.c_int, [C.c_char_p, u32, C.POINTER(u32p), C.POINTER(u32p)]),
            "vfr_con_from_incid": (C.c_int, [u32p, u32, u32, u32p, u32p, u32p, u32p]),
            "vfr_write_con": (C.c_int, [C.c_char_p, u32p, u32, u32]),
            "vfr_read_extf": (C.c_int, [C.c_char_p, u32, C.POINTER(u32p)]),
            "vfr_write_extf": (C.c_int, [C.c_char_p, u32p, u32]),
            "vfr_read_tf": (C.c_int, [C.c_char_p, f32p]),
            "vfr_write_tf": (C.c_int, [C.c_char_p, f32p]),
            "vfr_default_tf": (None, [f32p]),
            "vfr_read_lmt": (C.c_int, [C.c_char_p, f32p]),
            "vfr_write_lmt": (C.c_int, [C.c_char_p, f32p]),
            "vfr_build_lmt": (None, [f32p, u32, u32p, u32, f32p]),
            "vfr_write_incid": (C.c_int, [C.c_char_p, u32p, u32, u32]),
            "vfr_write_ppm": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint8), u32, u32]),
            "vfr_synth_grid": (C.c_int, [u32, C.c_uint64, u32, C.POINTER(f32p), u32p, C.POINTER(u32p), u32p]),
            "vfr_synth_tf": (None, [C.c_uint64, f32p]),
            "vfr_synth_rotate": (None, [f32p, u32, f32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != VFR_OK:
        raise VfrError(rc, lib().vfr_last_error().decode())


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


# ------------------------------------------------------------------ host helpers

def synth_grid(n: int, seed: int, variant: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Seeded hex-grid mesh: (xyzs float32 [nV,4], tets uint32 [nT,4])."""
    L = lib()
    xp, tp = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)()
    nv, nt = C.c_uint32(), C.c_uint32()
    _check(L.vfr_synth_grid(n, seed, variant, C.byref(xp), C.byref(nv), C.byref(tp), C.byref(nt)))
    try:
        x = np.ctypeslib.as_array(xp, shape=(nv.value * 4,)).copy().reshape(-1, 4)
        t = np.ctypeslib.as_array(tp, shape=(nt.value * 4,)).copy().reshape(-1, 4)
    finally:
        L.vfr_free(C.cast(xp, C.c_void_p))
        L.vfr_free(C.cast(tp, C.c_void_p))
    return x, t


def synth_tf(seed: int) -> np.ndarray:
    out = np.zeros(1024, np.float32)
    lib().vfr_synth_tf(seed, fptr(out))
    return out


def synth_rotate(xyzs: np.ndarray, d) -> np.ndarray:
    x = np.ascontiguousarray(xyzs, np.float32).copy()
    dd = np.asarray(d, np.float32)
    lib().vfr_synth_rotate(fptr(x), x.shape[0], fptr(dd))
    return x


def psi_generate() -> np.ndarray:
    out = np.zeros(512 * 512, np.float32)
    _check(lib().vfr_psi_generate(fptr(out)))
    return out.reshape(512, 512)


def normalize(xyzs: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(xyzs, np.float32).copy()
    lib().vfr_normalize_vertices(fptr(x), x.shape[0])
    return x


def build_lmt(xyzs: np.ndarray, tets: np.ndarray) -> np.ndarray:
    out = np.zeros(3, np.float32)
    x = np.ascontiguousarray(xyzs, np.float32)
    t = np.ascontiguousarray(tets, np.uint32)
    lib().vfr_build_lmt(fptr(x), x.shape[0], uptr(t), t.shape[0], fptr(out))
    return out


def default_tf() -> np.ndarray:
    out = np.zeros(1024, np.float32)
    lib().vfr_default_tf(fptr(out))
    return out


def write_off(path: str, xyzs: np.ndarray, tets: np.ndarray) -> None:
    x = np.ascontiguousarray(xyzs, np.float32)
    t = np.ascontiguousarray(tets, np.uint32)
    _check(lib().vfr_write_off(path.encode(), fptr(x), x.shape[0], uptr(t), t.shape[0]))


def write_tf(path: str, tf: np.ndarray) -> None:
    _check(lib().vfr_write_tf(path.encode(), fptr(np.ascontiguousarray(tf, np.float32))))


def gl_rotation(rotx_deg: float, roty_deg: float) -> np.ndarray:
    out = np.zeros(16, np.float32)
    lib().vfr_gl_rotation(C.c_float(rotx_deg), C.c_float(roty_deg), fptr(out))
    return out


def rotate_volume(xyzs: np.ndarray, mv) -> np.ndarray:
    x = np.ascontiguousarray(xyzs, np.float32).copy()
    m = np.ascontiguousarray(mv, np.float32).reshape(16)
    lib().vfr_rotate_volume(fptr(x), x.shape[0], fptr(m))
    return x


def read_incid(path: str, n_verts: int) -> tuple[np.ndarray, np.ndarray]:
    """The per-vertex tet lists of a .incid file: (offsets [n_verts + 1], ids)."""
    L = lib()
    op, ip = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint32)()
    _check(L.vfr_read_incid(path.encode(), n_verts, C.byref(op), C.byref(ip)))
    try:
        off = np.ctypeslib.as_array(op, shape=(n_verts + 1,)).copy()
        ids = np.ctypeslib.as_array(ip, shape=(max(1, int(off[-1])),)).copy()[: int(off[-1])]
    finally:
        L.vfr_free(C.cast(op, C.c_void_p))
        L.vfr_free(C.cast(ip, C.c_void_p))
    return off, ids


def con_from_incid(tets: np.ndarray, n_verts: int, offsets: np.ndarray, ids: np.ndarray) -> tuple[np.ndarray, int]:
    """The reference's buildCon over incidence lists (host): (con [nT, 4], n_ext)."""
    t = np.ascontiguousarray(tets, np.uint32)
    o = np.ascontiguousarray(offsets, np.uint32)
    i = np.ascontiguousarray(ids if ids.size else np.zeros(1), np.uint32)
    con = np.zeros((t.shape[0], 4), np.uint32)
    n = C.c_uint32()
    _check(lib().vfr_con_from_incid(uptr(t), t.shape[0], n_verts, uptr(o), uptr(i), uptr(con), C.byref(n)))
    return con, n.value


def read_tf(path: str) -> np.ndarray:
    out = np.zeros(1024, np.float32)
    _check(lib().vfr_read_tf(path.encode(), fptr(out)))
    return out


def to_rgb8(rgba: np.ndarray) -> np.ndarray:
    """uchar RGB the way the reference stores it: (uchar)(255*min(c,1))
    (vfRayKernelCPU.cc:546-552), computed in float32."""
    c = np.minimum(rgba[..., :3].astype(np.float32), np.float32(1.0))
    return (np.float32(255.0) * c).astype(np.int32).astype(np.uint8)


# ---------------------------------------------------------------------- views

def ortho_view(width: int, height: int, tf_id: int, matrix=None) -> View:
    v = View()
    v.width, v.height, v.projection, v.tf_id = width, height, PROJ_ORTHO, tf_id
    m = np.eye(3, 4, dtype=np.float32) if matrix is None else np.asarray(matrix, np.float32).reshape(3, 4)
    for i in range(12):
        v.ortho[i] = float(m.flat[i])
    return v


def perspective_view(width: int, height: int, tf_id: int, eye, target=(0, 0, 0), up=None,
                     fov_y_deg: float = 60.0) -> View:
    eye = np.asarray(eye, np.float64)
    target = np.asarray(target, np.float64)
    if up is None:
        f = target - eye
        f /= np.linalg.norm(f)
        up = np.array([0.0, 0.0, 1.0]) if abs(f[2]) < 0.9 else np.array([0.0, 1.0, 0.0])
    v = View()
    v.width, v.height, v.projection, v.tf_id = width, height, PROJ_PERSPECTIVE, tf_id
    for i in range(3):
        v.eye[i], v.target[i], v.up[i] = float(eye[i]), float(target[i]), float(up[i])
    v.fov_y_deg = fov_y_deg
    return v


def view_matrix(v: View) -> tuple[int, np.ndarray, float, float]:
    """(projection, the float32 3x4 matrix, k, kx) the library derives for a
    view (same double maths, rounded once), for the oracle.  Perspective: k =
    tan(fov_y/2) and kx = k W/H, the half-extents of the image plane at unit
    depth; orthographic: k = world length of a unit step along -Z (= kx)."""
    if v.projection == PROJ_ORTHO:
        M = np.array(v.ortho[:], np.float32)
        A = M.reshape(3, 4)[:, :3].astype(np.float64)
        w = np.linalg.inv(A)[:, 2]
        k = float(np.float32(np.sqrt((w * w).sum())))
        return 0, M, k, k
    E = np.array(v.eye[:], np.float64)
    f = np.array(v.target[:], np.float64) - E
    up = np.array(v.up[:], np.float64)
    f /= np.sqrt((f * f).sum())
    r = np.array([f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]])
    r /= np.sqrt((r * r).sum())
    u = np.array([r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]])
    k = np.tan(float(np.float32(v.fov_y_deg)) * np.pi / 360.0)
    kx = k * float(v.width) / float(v.height)
    M = np.zeros(12, np.float32)
    for i, (row, sc) in enumerate(((r, 1.0 / kx), (u, 1.0 / k), (f, 1.0))):
        for a in range(3):
            M[i * 4 + a] = np.float32(sc * row[a])
        M[i * 4 + 3] = np.float32(-sc * (row[0] * E[0] + row[1] * E[1] + row[2] * E[2]))
    return 1, M, float(np.float32(k)), float(np.float32(kx))


# -------------------------------------------------------------------- objects

class Context:
    def __init__(self, device: int = 0):
        L = lib()
        self._p = C.c_void_p()
        _check(L.vfr_ctx_create(device, C.byref(self._p)))

    def close(self) -> None:
        if self._p:
            lib().vfr_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def tf(self, rgba: np.ndarray) -> int:
        t = np.ascontiguousarray(rgba, np.float32).reshape(-1)
        i = C.c_int()
        _check(lib().vfr_tf_create(self._p, fptr(t), C.byref(i)))
        return i.value

    def bandwidth_probe(self, nbytes: int = 1 << 30, iters: int = 20) -> float:
        g = C.c_double()
        _check(lib().vfr_bandwidth_probe(self._p, nbytes, iters, C.byref(g)))
        return g.value

    def psi_load(self, table: np.ndarray | None) -> None:
        _check(lib().vfr_psi_load(self._p, None if table is None else fptr(np.ascontiguousarray(table, np.float32))))

    def psi_read(self) -> np.ndarray:
        out = np.zeros(512 * 512, np.float32)
        _check(lib().vfr_psi_read(self._p, fptr(out)))
        return out.reshape(512, 512)

    def mesh(self, xyzs, tets, con=None, extf=None, lmt=None, normalize=True, morton=True) -> "Mesh":
        return Mesh(self, xyzs, tets, con, extf, lmt, normalize, morton)

    def plan(self, mesh: "Mesh", views, mode=MODE_FTB, schedule=SCHED_BATCHED, term_alpha=0.99,
             use_graph=False, waves_per_cu=0, refill=0, persistent=False, occupancy=0,
             max_entries=0, pipelined=False, composite=COMPOSITE_PPI, psi_table=False, wide=False,
             wave_queue=False, lds_records=False, face_memo=False) -> "Plan":
        return Plan(self, mesh, views, mode, schedule, term_alpha, use_graph, waves_per_cu, refill,
                    persistent, occupancy, max_entries, pipelined, composite, psi_table, wide,
                    wave_queue, lds_records, face_memo)

    def render(self, mesh: "Mesh", views, **kw) -> tuple[list[dict], dict]:
        p = self.plan(mesh, views, **kw)
        try:
            st = p.execute()
            return [p.read(i) for i in range(len(p.views))], st
        finally:
            p.close()


class Mesh:
    def __init__(self, ctx: Context, xyzs, tets, con=None, extf=None, lmt=None, normalize=True, morton=True):
        self.ctx = ctx
        x = np.ascontiguousarray(xyzs, np.float32).reshape(-1)
        t = np.ascontiguousarray(tets, np.uint32).reshape(-1)
        c = None if con is None else np.ascontiguousarray(con, np.uint32).reshape(-1)
        e = None if extf is None else np.ascontiguousarray(extf, np.uint32).reshape(-1)
        lm = None if lmt is None else np.ascontiguousarray(lmt, np.float32)
        flags = (MESH_NORMALIZE if normalize else 0) | (MESH_MORTON if morton else 0)
        self._p = C.c_void_p()
        _check(lib().vfr_mesh_create(ctx._p, fptr(x), x.size // 4, uptr(t), t.size // 4, uptr(c), uptr(e),
                                     0 if e is None else e.size // 2, None if lm is None else fptr(lm),
                                     flags, C.byref(self._p)))
        inf = MeshInfo()
        _check(lib().vfr_mesh_get_info(self._p, C.byref(inf)))
        self.info = {k: getattr(inf, k) for k, _ in inf._fields_}

    def connectivity(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        con = np.zeros(self.info["n_tets"] * 4, np.uint32)
        extf = np.zeros(max(1, self.info["n_ext_faces"] * 2), np.uint32)
        lmt = np.zeros(3, np.float32)
        _check(lib().vfr_mesh_connectivity(self._p, uptr(con), uptr(extf), fptr(lmt)))
        return con.reshape(-1, 4), extf[: self.info["n_ext_faces"] * 2].reshape(-1, 2), lmt

    def rotate(self, mv, normalize=True) -> None:
        """rotateVolume + normalizeVertices on the resident vertices."""
        m = np.ascontiguousarray(mv, np.float32).reshape(16)
        _check(lib().vfr_mesh_rotate(self._p, fptr(m), MESH_NORMALIZE if normalize else 0))

    def vertices(self) -> np.ndarray:
        out = np.zeros(self.info["n_verts"] * 4, np.float32)
        _check(lib().vfr_mesh_vertices(self._p, fptr(out)))
        return out.reshape(-1, 4)

    def close(self) -> None:
        if self._p:
            lib().vfr_mesh_destroy(self._p)
            self._p = C.c_void_p()


class Plan:
    def __init__(self, ctx, mesh, views, mode, schedule, term_alpha, use_graph, waves_per_cu, refill,
                 persistent=False, occupancy=0, max_entries=0, pipelined=False, composite=COMPOSITE_PPI,
                 psi_table=False, wide=False, wave_queue=False, lds_records=False, face_memo=False):
        self.ctx, self.mesh = ctx, mesh
        self.views = list(views)
        arr = (View * len(self.views))(*self.views)
        flags = ((RENDER_PERSISTENT if persistent else 0) | (RENDER_PIPELINED if pipelined else 0)
                 | (RENDER_PSI_TABLE if psi_table else 0) | (RENDER_WIDE if wide else 0)
                 | (RENDER_WAVE_QUEUE if wave_queue else 0) | (RENDER_LDS_RECORDS if lds_records else 0)
                 | (RENDER_FACE_MEMO if face_memo else 0))
        o = RenderOpts(mode, schedule, term_alpha, waves_per_cu, refill, 1 if use_graph else 0, max_entries,
                       flags, occupancy, composite)
        self._p = C.c_void_p()
        _check(lib().vfr_plan_create(ctx._p, mesh._p, arr, len(self.views), C.byref(o), C.byref(self._p)))

    def autotune(self) -> dict:
        f, o, w, r, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
        _check(lib().vfr_plan_autotune(self._p, C.byref(f), C.byref(o), C.byref(w), C.byref(r), C.byref(ms)))
        return {"persistent": bool(f.value & RENDER_PERSISTENT), "pipelined": bool(f.value & RENDER_PIPELINED),
                "wave_queue": bool(f.value & RENDER_WAVE_QUEUE),
                "lds_records": bool(f.value & RENDER_LDS_RECORDS),
                "face_memo": bool(f.value & RENDER_FACE_MEMO),
                "occupancy": o.value, "waves_per_cu": w.value, "refill": r.value, "traversal_ms": ms.value}

    def execute(self) -> dict:
        st = Stats()
        _check(lib().vfr_plan_execute(self._p, C.byref(st)))
        return st.as_dict()

    def frame_times(self) -> np.ndarray:
        out = np.zeros(len(self.views), np.float64)
        _check(lib().vfr_plan_frame_times(self._p, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def read(self, i: int, rgba=True, entry=True, steps=True) -> dict:
        v = self.views[i]
        n = v.width * v.height
        out = {}
        fo = FrameOut()
        if rgba:
            out["rgba"] = np.zeros((v.height, v.width, 4), np.float32)
            fo.rgba = fptr(out["rgba"])
        if entry:
            out["entry_ext"] = np.zeros((v.height, v.width), np.uint32)
            out["entry_tet"] = np.zeros((v.height, v.width), np.uint32)
            fo.entry_ext = uptr(out["entry_ext"])
            fo.entry_tet = uptr(out["entry_tet"])
        if steps:
            out["steps"] = np.zeros((v.height, v.width), np.uint32)
            fo.steps = uptr(out["steps"])
        _check(lib().vfr_plan_read(self._p, i, C.byref(fo)))
        assert n == v.width * v.height
        return out

    def close(self) -> None:
        if self._p:
            lib().vfr_plan_destroy(self._p)
            self._p = C.c_void_p()


class Server:
    """Resident meshes + continuous batching (vfr_server_*)."""

    def __init__(self, ctx: Context, meshes, hbm_budget: int = 0, streams: int = 0, max_batch: int = 0,
                 term_alpha: float = 0.0, persistent: bool = False, occupancy: int = 0, waves_per_cu: int = 0,
                 refill: int = 0, psi_table: bool = False, wave_queue: bool = False,
                 composite: int = COMPOSITE_PPI, max_batch_pixels: int = 0, lds_records: bool = False,
                 schedule: int = 0):
        self.ctx = ctx
        self.meshes = list(meshes)
        arr = (C.c_void_p * len(self.meshes))(*[m._p.value for m in self.meshes])
        flags = ((RENDER_PERSISTENT if persistent else 0) | (RENDER_PSI_TABLE if psi_table else 0)
                 | (RENDER_WAVE_QUEUE if wave_queue else 0) | (RENDER_LDS_RECORDS if lds_records else 0))
        cfg = ServerCfg(hbm_budget, streams, max_batch, term_alpha, flags, occupancy, waves_per_cu, refill,
                        composite, max_batch_pixels, schedule)
        self._p = C.c_void_p()
        _check(lib().vfr_server_create(ctx._p, C.byref(cfg), arr, len(self.meshes), C.byref(self._p)))

    def submit(self, mesh: int, view: View, keep_image: bool = False) -> int:
        t = C.c_uint64()
        _check(lib().vfr_server_submit(self._p, C.byref(Request(mesh, 1 if keep_image else 0, view)),
                                       C.byref(t)))
        return t.value

    def poll(self, ticket: int) -> dict:
        r = Result()
        _check(lib().vfr_server_poll(self._p, ticket, C.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    def wait(self, ticket: int, image: bool = False):
        r = Result()
        img = None
        ptr = None
        if image:
            r0 = Result()
            _check(lib().vfr_server_poll(self._p, ticket, C.byref(r0)))
            img = np.zeros((r0.height, r0.width, 4), np.float32)
            ptr = fptr(img)
        _check(lib().vfr_server_wait(self._p, ticket, C.byref(r), ptr))
        out = {k: getattr(r, k) for k, _ in r._fields_}
        if image:
            out["rgba"] = img
        return out

    def replay(self, requests, offsets_s) -> dict:
        n = len(requests)
        arr = (Request * n)(*requests)
        off = np.ascontiguousarray(offsets_s, np.float64)
        rep = Report()
        _check(lib().vfr_server_replay(self._p, arr, off.ctypes.data_as(C.POINTER(C.c_double)), n,
                                       C.byref(rep)))
        return rep.as_dict()

    def close(self) -> None:
        if self._p:
            lib().vfr_server_destroy(self._p)
            self._p = C.c_void_p()
