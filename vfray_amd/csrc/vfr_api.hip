/*
 * vfr_api.hip -- the C ABI of libvfr.so (include/vfr.h): contexts, device
 * meshes, transfer functions, the psi table, render plans.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/vfr.h"
#include "vfr_internal.h"

#define CHECK_HIP(x)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      vfr_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));       \
      return VFR_EHIP;                                                                    \
    }                                                                                     \
  } while (0)

#define CHECK_RC(x)              \
  do {                           \
    int rc_ = (x);               \
    if (rc_ != VFR_OK) return rc_; \
  } while (0)

static const uint32_t kTfSlots = 4096;

struct vfr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  float* psi = nullptr;      /* [512*512] device */
  float4* tf = nullptr;      /* [kTfSlots*256] device */
  uint32_t n_tf = 0;
  int n_cu = 0;
};

struct vfr_mesh {
  vfr_ctx* ctx = nullptr;
  DevMesh d;
  std::vector<uint32_t> h_extf; /* caller (tet, face) pairs */
  float bb_lo[3] = {0, 0, 0}, bb_hi[3] = {0, 0, 0};
  uint32_t version = 0;         /* bumped by vfr_mesh_rotate */
  const uint4* tv4() const { return reinterpret_cast<const uint4*>(d.tv); }
  const uint4* fvid4() const { return reinterpret_cast<const uint4*>(d.fvid); }
};

extern "C" const char* vfr_version(void) { return "vfr 0.1 gfx950"; }

/* Accessors for the server (vfr_server.hip). */
int vfr_ctx_device(const vfr_ctx* c) { return c->device; }
void* vfr_ctx_stream(const vfr_ctx* c) { return c->stream; }
const uint32_t* vfr_mesh_host_extf(const vfr_mesh* m) { return m->h_extf.data(); }
int vfr_ctx_cus(const vfr_ctx* c) { return c->n_cu; }
const float4* vfr_ctx_tf(const vfr_ctx* c) { return c->tf; }
uint32_t vfr_ctx_n_tf(const vfr_ctx* c) { return c->n_tf; }
const float* vfr_ctx_psi(const vfr_ctx* c) { return c->psi; }
const vfr_ctx* vfr_mesh_ctx(const vfr_mesh* m) { return m->ctx; }
const DevMesh* vfr_mesh_dev(const vfr_mesh* m) { return &m->d; }
void vfr_mesh_bbox(const vfr_mesh* m, float lo[3], float hi[3]) {
  for (int a = 0; a < 3; ++a) { lo[a] = m->bb_lo[a]; hi[a] = m->bb_hi[a]; }
}

/* ----------------------------------------------------------------- context */

extern "C" int vfr_ctx_create(int device, vfr_ctx** out) {
  if (!out) { vfr_set_error("vfr_ctx_create: null out"); return VFR_EINVAL; }
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
    vfr_set_error("no HIP device %d (found %d)", device, n);
    return VFR_ENODEV;
  }
  CHECK_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  CHECK_HIP(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    vfr_set_error("device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
    return VFR_ENODEV;
  }
  vfr_ctx* c = new (std::nothrow) vfr_ctx;
  if (!c) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&c->psi, sizeof(float) * 512 * 512) != hipSuccess ||
      hipMalloc((void**)&c->tf, sizeof(float4) * 256 * kTfSlots) != hipSuccess) {
    vfr_set_error("device allocation failed");
    vfr_ctx_destroy(c);
    return VFR_ENOMEM;
  }
  int rc = vfr_psi_load(c, nullptr);
  if (rc == VFR_OK) rc = vfr_noalloc_init(device); /* device is current */
  if (rc != VFR_OK) { vfr_ctx_destroy(c); return rc; }
  *out = c;
  return VFR_OK;
}

extern "C" void vfr_ctx_destroy(vfr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->psi) (void)hipFree(c->psi);
  if (c->tf) (void)hipFree(c->tf);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static std::once_flag g_psi_once;
static std::vector<float> g_psi;

static const float* closed_form_psi() {
  std::call_once(g_psi_once, [] {
    g_psi.resize(512 * 512);
    vfr_psi_generate(g_psi.data());
  });
  return g_psi.data();
}

extern "C" int vfr_psi_load(vfr_ctx* c, const float* table) {
  if (!c) { vfr_set_error("vfr_psi_load: null ctx"); return VFR_EINVAL; }
  const float* src = table ? table : closed_form_psi();
  CHECK_HIP(hipSetDevice(c->device));
  CHECK_HIP(hipMemcpy(c->psi, src, sizeof(float) * 512 * 512, hipMemcpyHostToDevice));
  return VFR_OK;
}

extern "C" int vfr_psi_read(vfr_ctx* c, float* out) {
  if (!c || !out) { vfr_set_error("vfr_psi_read: null argument"); return VFR_EINVAL; }
  CHECK_HIP(hipSetDevice(c->device));
  CHECK_HIP(hipMemcpy(out, c->psi, sizeof(float) * 512 * 512, hipMemcpyDeviceToHost));
  return VFR_OK;
}

extern "C" int vfr_tf_create(vfr_ctx* c, const float rgba[1024], int* tf_id) {
  if (!c || !rgba || !tf_id) { vfr_set_error("vfr_tf_create: null argument"); return VFR_EINVAL; }
  if (c->n_tf >= kTfSlots) { vfr_set_error("transfer-function table full (%u)", kTfSlots); return VFR_ENOMEM; }
  for (int i = 0; i < 1024; ++i)
    if (!(rgba[i] >= 0.0f && rgba[i] <= 1.0f)) {
      vfr_set_error("transfer function entry %d component %d = %g is outside [0,1]", i / 4, i % 4,
                    (double)rgba[i]);
      return VFR_EINVAL;
    }
  CHECK_HIP(hipSetDevice(c->device));
  CHECK_HIP(hipMemcpy(c->tf + (size_t)c->n_tf * 256, rgba, sizeof(float) * 1024, hipMemcpyHostToDevice));
  *tf_id = (int)c->n_tf++;
  return VFR_OK;
}

extern "C" int vfr_bandwidth_probe(vfr_ctx* c, uint64_t bytes, int iters, double* gbps) {
  if (!c || !gbps || bytes < 16 || iters < 1) { vfr_set_error("vfr_bandwidth_probe: bad argument"); return VFR_EINVAL; }
  CHECK_HIP(hipSetDevice(c->device));
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    if (a) (void)hipFree(a);
    vfr_set_error("probe allocation of 2 x %llu bytes failed", (unsigned long long)bytes);
    return VFR_ENOMEM;
  }
  int rc = vfr_k_clear_u32((uint32_t*)a, bytes / 4, c->stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == VFR_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = VFR_EHIP;
  if (rc == VFR_OK) rc = vfr_k_copy(a, b, bytes, c->stream);
  if (rc == VFR_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = VFR_EHIP;
  for (int i = 0; rc == VFR_OK && i < iters; ++i) rc = vfr_k_copy(a, b, bytes, c->stream);
  float ms = 0;
  if (rc == VFR_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = VFR_EHIP;
  if (rc == VFR_OK) *gbps = 2.0 * (double)(bytes / 16 * 16) * iters / (ms * 1e-3) / 1e9;
  else if (rc == VFR_EHIP) vfr_set_error("bandwidth probe failed");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

/* -------------------------------------------------------------------- mesh */

extern "C" int vfr_mesh_create(vfr_ctx* c, const float* xyzs, uint32_t nV, const uint32_t* tets,
                               uint32_t nT, const uint32_t* con, const uint32_t* extf,
                               uint32_t n_ext, const float* lmt, uint32_t flags, vfr_mesh** out) {
  if (!c || !xyzs || !tets || !out || nV < 4 || nT < 1) {
    vfr_set_error("vfr_mesh_create: bad argument");
    return VFR_EINVAL;
  }
  if (nT >= VFR_MAX_TETS) {
    /* neighbour words hold next << 2 */
    vfr_set_error("vfr_mesh_create: %u tets exceeds the limit of %u", nT, VFR_MAX_TETS - 1u);
    return VFR_EINVAL;
  }
  if (nV >= VFR_MAX_VERTS) {
    /* tet records hold vertex ids in VFR_REC_VERT_BITS bits */
    vfr_set_error("vfr_mesh_create: %u vertices exceeds the limit of %u", nV, VFR_MAX_VERTS - 1u);
    return VFR_EINVAL;
  }
  if (extf && !con) { vfr_set_error("vfr_mesh_create: extf given without con"); return VFR_EINVAL; }
  *out = nullptr;
  CHECK_HIP(hipSetDevice(c->device));
  auto t0 = std::chrono::steady_clock::now();
  vfr_mesh* m = new (std::nothrow) vfr_mesh;
  if (!m) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  m->ctx = c;
  m->d.n_verts = nV;
  m->d.n_tets = nT;
  /* ids are checked, vertices normalized and limits built on the device */
  int rc = vfr_dev_build_mesh(m->d, xyzs, tets, con, extf, n_ext, (flags & VFR_MESH_NORMALIZE) != 0,
                              (flags & VFR_MESH_MORTON) != 0, lmt, m->bb_lo, m->bb_hi, c->stream);
  if (rc != VFR_OK) { vfr_mesh_destroy(m); return rc; }
  if (!(m->d.lmt[0] > 0.0f)) { vfr_set_error("max edge length %g is not positive", (double)m->d.lmt[0]); vfr_mesh_destroy(m); return VFR_EMESH; }
  m->h_extf.resize(2 * (size_t)m->d.n_ext);
  if (m->d.n_ext &&
      hipMemcpy(m->h_extf.data(), m->d.extf, sizeof(uint32_t) * 2 * m->d.n_ext, hipMemcpyDeviceToHost) != hipSuccess) {
    vfr_set_error("copy of external faces failed");
    vfr_mesh_destroy(m);
    return VFR_EHIP;
  }
  m->d.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = m;
  return VFR_OK;
}

extern "C" void vfr_mesh_destroy(vfr_mesh* m) {
  if (!m) return;
  if (m->ctx) (void)hipSetDevice(m->ctx->device);
  void* ps[] = {m->d.rec, m->d.tv, m->d.vert, m->d.ext_int, m->d.fvid, m->d.extf, m->d.perm, m->d.vperm,
                m->d.con};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete m;
}

extern "C" int vfr_mesh_get_info(const vfr_mesh* m, vfr_mesh_info* out) {
  if (!m || !out) { vfr_set_error("vfr_mesh_get_info: null argument"); return VFR_EINVAL; }
  out->n_verts = m->d.n_verts;
  out->n_tets = m->d.n_tets;
  out->n_ext_faces = m->d.n_ext;
  out->max_edge_length = m->d.lmt[0];
  out->max_z = m->d.lmt[1];
  out->min_z = m->d.lmt[2];
  out->device_bytes = m->d.bytes;
  out->build_ms = m->d.build_ms;
  return VFR_OK;
}

extern "C" int vfr_mesh_connectivity(const vfr_mesh* m, uint32_t* con_out, uint32_t* extf_out,
                                     float* lmt_out) {
  if (!m) { vfr_set_error("vfr_mesh_connectivity: null mesh"); return VFR_EINVAL; }
  CHECK_HIP(hipSetDevice(m->ctx->device));
  if (con_out)
    CHECK_HIP(hipMemcpy(con_out, m->d.con, sizeof(uint32_t) * 4 * (size_t)m->d.n_tets, hipMemcpyDeviceToHost));
  if (extf_out && m->d.n_ext) memcpy(extf_out, m->h_extf.data(), sizeof(uint32_t) * 2 * m->d.n_ext);
  if (lmt_out) memcpy(lmt_out, m->d.lmt, sizeof(float) * 3);
  return VFR_OK;
}

extern "C" int vfr_mesh_rotate(vfr_mesh* m, const float mv[16], uint32_t flags) {
  if (!m || !mv) { vfr_set_error("vfr_mesh_rotate: null argument"); return VFR_EINVAL; }
  if (flags & ~VFR_MESH_NORMALIZE) { vfr_set_error("vfr_mesh_rotate: unknown flags 0x%x", flags); return VFR_EINVAL; }
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(mv[k])) { vfr_set_error("vfr_mesh_rotate: matrix entry %d is not finite", k); return VFR_EINVAL; }
  CHECK_HIP(hipSetDevice(m->ctx->device));
  float lo[3], hi[3];
  CHECK_RC(vfr_k_rotate_normalize(reinterpret_cast<float4*>(m->d.vert), m->d.n_verts, mv,
                                  (flags & VFR_MESH_NORMALIZE) ? 1 : 0, lo, hi, m->ctx->stream));
  for (int a = 0; a < 3; ++a) { m->bb_lo[a] = lo[a]; m->bb_hi[a] = hi[a]; }
  ++m->version;
  return VFR_OK;
}

extern "C" int vfr_mesh_vertices(const vfr_mesh* m, float* out) {
  if (!m || !out) { vfr_set_error("vfr_mesh_vertices: null argument"); return VFR_EINVAL; }
  CHECK_HIP(hipSetDevice(m->ctx->device));
  std::vector<float> v(4 * (size_t)m->d.n_verts);
  std::vector<uint32_t> vp(m->d.n_verts);
  CHECK_HIP(hipMemcpy(v.data(), m->d.vert, sizeof(float) * v.size(), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(vp.data(), m->d.vperm, sizeof(uint32_t) * vp.size(), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < m->d.n_verts; ++i) memcpy(out + (size_t)vp[i] * 4, &v[(size_t)i * 4], 16);
  return VFR_OK;
}

/* -------------------------------------------------------------------- plan */

struct vfr_plan {
  vfr_ctx* ctx = nullptr;
  const vfr_mesh* mesh = nullptr;
  vfr_render_opts opts{};
  uint32_t mesh_version = 0;              /* the mesh's version at creation */
  std::vector<DevView> hv;
  uint32_t n_views = 0;
  uint64_t n_pix = 0;
  uint32_t n_tiles = 0;
  uint32_t trav_proj = VFR_TRAV_MIXED;    /* projections of the plan's views */
  uint32_t tf_lo = 0, tf_n = 0;           /* TF slots the views use, if few  */
  DevView* views = nullptr;
  uint32_t* owner = nullptr;   /* REF owner or FTB count, [n_pix] */
  uint64_t* entries = nullptr; /* FTB, [n_pix * kmax] */
  uint32_t kmax = VFR_K;       /* FTB entry slots per pixel */
  uint64_t* rays = nullptr;    /* [n_pix] */
  float4* pverts = nullptr;    /* [n_views * n_verts] P-space vertices */
  float4* memo = nullptr;      /* REF face-plane memo, [n_views * n_tets * 4 * 2] */
  uint32_t* block_off = nullptr; /* batched compaction: per-block ray counts, then offsets */
  uint64_t* tile_mask = nullptr;  /* ... each tile's active-pixel mask, then its view */
  float4* rgba = nullptr;
  uint32_t* steps = nullptr;
  uint32_t* entry = nullptr;
  uint32_t* counters = nullptr;           /* [n_views + 2]: per-view ray counts,
                                             batched ray count, work-queue head */
  unsigned long long* stats = nullptr;    /* [ST_N] */
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> fev;            /* per-view frame events */
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool executed = false;
  std::vector<double> frame_ms;           /* last PER_VIEW execute, per view */
  /* pipelined batched execution: view chunks whose preprocessing (stream 1)
   * overlaps the walk of the chunks before them (stream 2) */
  struct Chunk {
    uint32_t v0, nv, tile0, ntiles, proj;
    uint64_t pix0, npix;
  };
  std::vector<Chunk> chunks;
  hipStream_t st2 = nullptr;
  std::vector<hipEvent_t> cev;            /* per chunk: ready, walk start, walk end; + join */
};

/* The REF walk reads the face-plane memo (VFR_RENDER_FACE_MEMO). */
static bool plan_memo_on(const vfr_plan* p) {
  return p->opts.mode == VFR_MODE_REF && (p->opts.flags & VFR_RENDER_FACE_MEMO) && p->memo;
}

/* Device bytes of a plan's face-plane memo: 2 float4 per tet face per view. */
static size_t plan_memo_bytes(const vfr_plan* p) {
  return (size_t)p->n_views * p->mesh->d.n_tets * 4u * 2u * sizeof(float4);
}


/* View set-up on the host, in double, rounded once to float.  bb_lo/bb_hi:
 * the mesh's bounding box (every vertex must lie in front of a pinhole eye). */
int vfr_make_view(const vfr_view& in, const float* bb_lo, const float* bb_hi, DevView* o) {
  memset(o, 0, sizeof(*o));
  if (in.width < 2 || in.height < 2 || in.width > 16384 || in.height > 16384) {
    vfr_set_error("image size %ux%u outside [2,16384]", in.width, in.height);
    return VFR_EINVAL;
  }
  o->W = in.width;
  o->H = in.height;
  o->proj = in.projection;
  o->tf = (uint32_t)in.tf_id;
  if (in.projection == VFR_PROJ_ORTHO) {
    double A[3][3];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 4; ++c) o->M[r * 4 + c] = in.ortho[r * 4 + c];
      for (int c = 0; c < 3; ++c) A[r][c] = in.ortho[r * 4 + c];
    }
    double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                 A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                 A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (!(fabs(det) > 1e-12)) { vfr_set_error("ortho transform is singular"); return VFR_EINVAL; }
    /* world direction of a unit step along -Z: third column of A^-1 */
    double w0 = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    double w1 = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]) / det;
    double w2 = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    o->k = (float)sqrt(w0 * w0 + w1 * w1 + w2 * w2);
  } else if (in.projection == VFR_PROJ_PERSPECTIVE) {
    if (!(in.fov_y_deg > 1.0f && in.fov_y_deg < 170.0f)) {
      vfr_set_error("fov %g outside (1,170) degrees", (double)in.fov_y_deg);
      return VFR_EINVAL;
    }
    double E[3], f[3], up[3];
    for (int a = 0; a < 3; ++a) { E[a] = in.eye[a]; f[a] = (double)in.target[a] - in.eye[a]; up[a] = in.up[a]; }
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0)) { vfr_set_error("eye equals target"); return VFR_EINVAL; }
    for (int a = 0; a < 3; ++a) f[a] /= fl;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(rl > 1e-9)) { vfr_set_error("up vector parallel to the view direction"); return VFR_EINVAL; }
    for (int a = 0; a < 3; ++a) r[a] /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    /* fov_y spans the image height; the width spans tan(fov_y/2) W/H, so
     * pixels are square whatever the aspect ratio */
    double k = tan(in.fov_y_deg * M_PI / 360.0), kx = k * (double)in.width / (double)in.height;
    const double* rows[3] = {r, u, f};
    const double sc[3] = {1.0 / kx, 1.0 / k, 1.0};
    for (int i = 0; i < 3; ++i) {
      for (int a = 0; a < 3; ++a) o->M[i * 4 + a] = (float)(sc[i] * rows[i][a]);
      o->M[i * 4 + 3] = (float)(-sc[i] * (rows[i][0] * E[0] + rows[i][1] * E[1] + rows[i][2] * E[2]));
    }
    o->k = (float)k;
    o->kx = (float)kx;
    /* every vertex must lie in front of the eye: check the bbox corners */
    for (int q = 0; q < 8; ++q) {
      double p[3] = {(q & 1) ? bb_hi[0] : bb_lo[0], (q & 2) ? bb_hi[1] : bb_lo[1],
                     (q & 4) ? bb_hi[2] : bb_lo[2]};
      double dc = f[0] * (p[0] - E[0]) + f[1] * (p[1] - E[1]) + f[2] * (p[2] - E[2]);
      if (!(dc > 1e-3)) {
        vfr_set_error("eye is inside or beside the mesh bounding box (depth %g)", dc);
        return VFR_EINVAL;
      }
    }
  } else {
    vfr_set_error("unknown projection %u", in.projection);
    return VFR_EINVAL;
  }
  return VFR_OK;
}

extern "C" void vfr_plan_destroy(vfr_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  void* ps[] = {p->views, p->owner, p->entries, p->rays, p->pverts, p->rgba, p->steps, p->entry,
                p->counters, p->stats, p->memo, p->block_off, p->tile_mask};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  for (auto& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : p->fev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : p->cev)
    if (e) (void)hipEventDestroy(e);
  if (p->st2) {
    (void)hipStreamSynchronize(p->st2);
    (void)hipStreamDestroy(p->st2);
  }
  delete p;
}

/* Largest number of entry faces any pixel of the plan has (device counting
 * pass: transform + raster without storing keys). */
static int plan_count_entries(vfr_plan* p, uint32_t* out) {
  hipStream_t st = p->ctx->stream;
  const vfr_mesh* m = p->mesh;
  CHECK_RC(vfr_k_clear_u32(p->owner, p->n_pix, st));
  CHECK_RC(vfr_k_clear_u32(p->counters, 1, st));
  CHECK_RC(vfr_k_xform_views(p->views, 0, p->n_views, m->d.n_verts, (const float4*)m->d.vert, p->pverts, st));
  CHECK_RC(vfr_k_raster_ftb(p->views, 0, p->n_views, m->d.n_ext, m->fvid4(), p->pverts, p->owner,
                            nullptr, 0, 0, p->stats, nullptr, st));
  CHECK_RC(vfr_k_max_u32(p->owner, p->n_pix, p->counters, st));
  CHECK_HIP(hipMemcpyAsync(out, p->counters, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipStreamSynchronize(st));
  return VFR_OK;
}

static int enqueue_batched(vfr_plan* p) {
  hipStream_t st = p->ctx->stream;
  const vfr_mesh* m = p->mesh;
  CHECK_RC(vfr_k_clear_u32(p->owner, p->n_pix, st));
  CHECK_RC(vfr_k_clear_u32(p->counters, p->n_views + 2, st));
  CHECK_RC(vfr_k_clear_u32((uint32_t*)p->stats, 2 * ST_N, st));
  CHECK_RC(vfr_k_xform_views(p->views, 0, p->n_views, m->d.n_verts, (const float4*)m->d.vert, p->pverts, st));
  if (plan_memo_on(p))
    CHECK_RC(vfr_k_face_memo(p->views, 0, p->n_views, m->d.n_tets, m->tv4(), p->pverts, p->memo, st));
  if (p->opts.mode == VFR_MODE_REF)
    CHECK_RC(vfr_k_raster_ref(p->views, 0, p->n_views, m->d.n_ext, m->fvid4(), p->pverts, p->owner, st));
  else
    CHECK_RC(vfr_k_raster_ftb(p->views, 0, p->n_views, m->d.n_ext, m->fvid4(), p->pverts, p->owner,
                              p->entries, p->kmax, p->n_pix, p->stats, nullptr, st));
  CHECK_RC(vfr_k_compact(p->views, p->n_views, 0, p->n_tiles, p->opts.mode, p->owner, p->owner,
                         p->entries, p->kmax, p->n_pix, p->rgba, p->steps, p->entry, p->rays, p->counters + p->n_views, st,
                         0, p->block_off, p->tile_mask));
  return VFR_OK;
}

/* Records the walk must address: the mesh's, or with VFR_RENDER_WIDE as many
 * as select the 64-bit-address walk variant. */
static uint64_t plan_records(const vfr_plan* p) {
  return (p->opts.flags & VFR_RENDER_WIDE) ? VFR_BIG_RECORDS : (uint64_t)p->mesh->d.n_tets;
}

static TravArgs trav_args(vfr_plan* p, uint64_t* rays, uint32_t* n_rays) {
  TravArgs a;
  memset(&a, 0, sizeof(a));
  const vfr_mesh* m = p->mesh;
  a.rec = m->d.rec;
  a.tv = m->tv4();
  a.vert = p->pverts;
  a.ext_int = m->d.ext_int;
  a.fvid = m->fvid4();
  a.tf = p->ctx->tf;
  a.psi = p->ctx->psi;
  a.views = p->views;
  a.rays = rays;
  a.n_rays = n_rays;
  a.owner = p->owner;
  a.cnt = p->owner;
  a.entries = p->entries;
  a.estride = p->n_pix;
  a.kmax = p->kmax;
  a.rgba = p->rgba;
  a.steps = p->steps;
  a.entry_out = p->entry;
  a.stats = p->stats;
  a.queue = p->counters + p->n_views + 1;
  a.term = p->opts.term_alpha;
  a.refill = p->opts.refill;
  a.tf_lds_base = p->tf_lo;
  a.tf_lds_n = p->tf_n;
  a.composite = p->opts.composite;
  a.psi_table = (p->opts.flags & VFR_RENDER_PSI_TABLE) ? 1u : 0u;
  a.lds_records = (p->opts.flags & VFR_RENDER_LDS_RECORDS) ? 1u : 0u;
  a.memo = plan_memo_on(p) ? p->memo : nullptr;
  a.memo_tets = p->mesh->d.n_tets;
  /* more TF slots than LDS holds for the launch: each block stages its own */
  a.tf_block = p->tf_n == 0 ? 1u : 0u;
  return a;
}

/* The walk kernel a batched front-to-back plan asks for. */
static int plan_walk_kernel(const vfr_plan* p) {
  if (p->opts.mode != VFR_MODE_FTB) return VFR_WALK_THREAD;
  if (p->opts.flags & VFR_RENDER_PERSISTENT) return VFR_WALK_REFILL;
  if (p->opts.flags & VFR_RENDER_WAVE_QUEUE) return VFR_WALK_QUEUE;
  return VFR_WALK_THREAD;
}

/* The batched traversal: one ray per thread over the rays of all views, or
 * a work-queue kernel (FTB, when asked; the queue head was cleared with the
 * plan's counters). */
static int launch_batched_traverse(vfr_plan* p) {
  TravArgs a = trav_args(p, p->rays, p->counters + p->n_views);
  const uint32_t waves = p->opts.waves_per_cu * (uint32_t)p->ctx->n_cu;
  return vfr_k_traverse(p->opts.mode, p->trav_proj, plan_records(p), p->opts.occupancy, a, (uint32_t)p->n_pix,
                        plan_walk_kernel(p), waves, p->ctx->stream);
}

extern "C" int vfr_plan_create(vfr_ctx* c, const vfr_mesh* m, const vfr_view* views, int n_views,
                               const vfr_render_opts* opts, vfr_plan** out) {
  if (!c || !m || !views || n_views < 1 || !out) { vfr_set_error("vfr_plan_create: bad argument"); return VFR_EINVAL; }
  if (m->ctx != c) { vfr_set_error("mesh belongs to another context"); return VFR_EINVAL; }
  *out = nullptr;
  vfr_render_opts o{};
  if (opts) o = *opts;
  if (o.mode > VFR_MODE_FTB || o.schedule > VFR_SCHED_BATCHED) { vfr_set_error("bad mode/schedule"); return VFR_EINVAL; }
  if (!(o.term_alpha > 0.0f)) o.term_alpha = 0.99f;
  if (o.term_alpha > 1.0f) { vfr_set_error("term_alpha %g > 1", (double)o.term_alpha); return VFR_EINVAL; }
  if (o.max_entries != 0 && o.max_entries != 8 && o.max_entries != 16 && o.max_entries != VFR_K_MAX) {
    vfr_set_error("max_entries must be 0 (chosen per plan), 8, 16 or %u", VFR_K_MAX);
    return VFR_EINVAL;
  }
  if (o.waves_per_cu == 0) o.waves_per_cu = 16;
  if (o.refill == 0) o.refill = 32;
  if (o.composite > VFR_COMPOSITE_LINEAR_OPACITY) {
    vfr_set_error("unknown compositing %u", o.composite);
    return VFR_EINVAL;
  }
  if (o.waves_per_cu > 64 || o.refill > 64) {
    vfr_set_error("waves_per_cu %u / refill %u out of range (<= 64)", o.waves_per_cu, o.refill);
    return VFR_EINVAL;
  }
  if ((o.flags & VFR_RENDER_PERSISTENT) && (o.flags & VFR_RENDER_WAVE_QUEUE)) {
    vfr_set_error("VFR_RENDER_PERSISTENT and VFR_RENDER_WAVE_QUEUE exclude each other");
    return VFR_EINVAL;
  }
  vfr_plan* p = new (std::nothrow) vfr_plan;
  if (!p) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  p->ctx = c;
  p->mesh = m;
  p->mesh_version = m->version;
  p->opts = o;
  p->n_views = (uint32_t)n_views;
  p->hv.resize(n_views);
  uint64_t pix = 0, tiles = 0;
  for (int i = 0; i < n_views; ++i) {
    int rc = vfr_make_view(views[i], m->bb_lo, m->bb_hi, &p->hv[i]);
    if (rc == VFR_OK && (views[i].tf_id < 0 || (uint32_t)views[i].tf_id >= c->n_tf)) {
      vfr_set_error("view %d: unknown transfer function %d", i, views[i].tf_id);
      rc = VFR_EINVAL;
    }
    if (rc == VFR_OK && o.mode == VFR_MODE_REF && views[i].projection != VFR_PROJ_ORTHO) {
      vfr_set_error("view %d: REF mode renders orthographic views only", i);
      rc = VFR_EINVAL;
    }
    if (rc != VFR_OK) { delete p; return rc; }
    p->hv[i].pix_off = (uint32_t)pix;
    p->hv[i].tile_off = (uint32_t)tiles;
    p->hv[i].tiles_x = (views[i].width + 7) / 8;
    p->hv[i].vert_off = (uint32_t)((uint64_t)i * m->d.n_verts);
    p->hv[i].rec_off = 0;
    p->hv[i].ext_off = 0;
    p->hv[i].wvert_off = 0;
    p->hv[i].n_ext = m->d.n_ext;
    p->hv[i].n_verts = m->d.n_verts;
    p->hv[i].max_edge = m->d.lmt[0];
    p->hv[i].inv_me = 1.0f / m->d.lmt[0];
    pix += (uint64_t)views[i].width * views[i].height;
    tiles += (uint64_t)p->hv[i].tiles_x * ((views[i].height + 7) / 8);
    if (pix >= (1ull << 31) || tiles >= (1ull << 31)) {
      vfr_set_error("plan exceeds 2^31 pixels");
      delete p;
      return VFR_EINVAL;
    }
  }
  p->n_pix = pix;
  p->n_tiles = (uint32_t)tiles;
  {
    uint32_t persp = 0;
    for (int i = 0; i < n_views; ++i) persp += views[i].projection == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
    p->trav_proj = persp == 0 ? VFR_TRAV_ORTHO : (persp == (uint32_t)n_views ? VFR_TRAV_PERSP : VFR_TRAV_MIXED);
    uint32_t lo = UINT32_MAX, hi = 0;
    for (int i = 0; i < n_views; ++i) {
      lo = std::min(lo, (uint32_t)views[i].tf_id);
      hi = std::max(hi, (uint32_t)views[i].tf_id);
    }
    /* few distinct slots: the walk reads them from LDS */
    if (hi - lo < VFR_TF_LDS_SLOTS) { p->tf_lo = lo; p->tf_n = hi - lo + 1; }
  }
  if ((uint64_t)n_views * m->d.n_verts >= VFR_MAX_VIEW_VERTS || (uint64_t)n_views * m->d.n_ext >= (1ull << 32)) {
    /* P-space vertices are addressed with 32-bit byte offsets */
    vfr_set_error("plan needs %llu view vertices (limit %llu): split the views over several plans",
                  (unsigned long long)n_views * m->d.n_verts, (unsigned long long)VFR_MAX_VIEW_VERTS - 1);
    delete p;
    return VFR_EINVAL;
  }
  int rc = VFR_OK;
  (void)hipSetDevice(c->device);
#define PALLOC(ptr, bytes)                                                        \
  if (rc == VFR_OK && hipMalloc((void**)&(ptr), (bytes) ? (bytes) : 16) != hipSuccess) { \
    vfr_set_error("plan allocation of %zu bytes failed", (size_t)(bytes));       \
    rc = VFR_ENOMEM;                                                              \
  }
  PALLOC(p->views, sizeof(DevView) * n_views);
  PALLOC(p->owner, sizeof(uint32_t) * pix);
  PALLOC(p->rays, sizeof(uint64_t) * pix);
  PALLOC(p->pverts, sizeof(float4) * (size_t)n_views * m->d.n_verts);
  PALLOC(p->rgba, sizeof(float4) * pix);
  PALLOC(p->steps, sizeof(uint32_t) * pix);
  PALLOC(p->entry, sizeof(uint32_t) * pix);
  PALLOC(p->counters, sizeof(uint32_t) * (n_views + 2 + VFR_PIPE_CHUNKS));
  PALLOC(p->stats, sizeof(unsigned long long) * ST_N);
  PALLOC(p->block_off, sizeof(uint32_t) * vfr_compact_blocks((uint32_t)tiles));
  PALLOC(p->tile_mask, sizeof(uint64_t) * tiles + sizeof(uint32_t) * tiles);
#undef PALLOC
  if (rc == VFR_OK && hipMemcpy(p->views, p->hv.data(), sizeof(DevView) * n_views, hipMemcpyHostToDevice) != hipSuccess) {
    vfr_set_error("view upload failed");
    rc = VFR_EHIP;
  }
  if (rc == VFR_OK && o.mode == VFR_MODE_FTB) {
    /* entry slots per pixel: as asked, or the fewest of 8/16/32 that hold
     * every pixel's entry faces, from a counting pass over all views */
    uint32_t k = o.max_entries;
    if (k == 0) {
      uint32_t mx = 0;
      rc = plan_count_entries(p, &mx);
      if (rc == VFR_OK && mx > VFR_K_MAX) {
        /* a ray along a boundary edge enters every face on it */
        vfr_set_error("a pixel has %u entry faces, more than the %u entry slots a plan holds; they would be "
                      "dropped", mx, VFR_K_MAX);
        rc = VFR_EMESH;
      }
      k = mx <= 8 ? 8u : (mx <= 16 ? 16u : VFR_K_MAX);
    }
    p->kmax = k;
    p->opts.max_entries = k;
    if (rc == VFR_OK && hipMalloc((void**)&p->entries, sizeof(uint64_t) * k * pix) != hipSuccess) {
      vfr_set_error("plan allocation of %zu bytes failed", (size_t)(sizeof(uint64_t) * k * pix));
      rc = VFR_ENOMEM;
    }
  }
  if (rc == VFR_OK && o.mode == VFR_MODE_REF && (o.flags & VFR_RENDER_FACE_MEMO) &&
      hipMalloc((void**)&p->memo, plan_memo_bytes(p)) != hipSuccess) {
    vfr_set_error("face-plane memo allocation of %zu bytes failed", plan_memo_bytes(p));
    rc = VFR_ENOMEM;
  }
  for (int i = 0; rc == VFR_OK && i < 3; ++i)
    if (hipEventCreate(&p->ev[i]) != hipSuccess) { vfr_set_error("event creation failed"); rc = VFR_EHIP; }
  if (rc == VFR_OK && o.schedule == VFR_SCHED_PER_VIEW) {
    p->fev.assign(2 * (size_t)n_views, nullptr);
    for (auto& e : p->fev)
      if (rc == VFR_OK && hipEventCreate(&e) != hipSuccess) { vfr_set_error("event creation failed"); rc = VFR_EHIP; }
  }
  if (rc == VFR_OK && o.schedule == VFR_SCHED_BATCHED && o.mode == VFR_MODE_FTB && !o.use_graph &&
      n_views >= 2) {
    /* chunks of consecutive views with about equal pixel counts */
    const uint32_t nc = std::min<uint32_t>(VFR_PIPE_CHUNKS, (uint32_t)n_views);
    uint32_t v = 0;
    uint64_t done = 0;
    for (uint32_t ci = 0; ci < nc && v < (uint32_t)n_views; ++ci) {
      vfr_plan::Chunk ch{};
      ch.v0 = v;
      ch.pix0 = p->hv[v].pix_off;
      ch.tile0 = p->hv[v].tile_off;
      const uint64_t goal = (pix * (ci + 1)) / nc;
      uint32_t persp = 0;
      while (v < (uint32_t)n_views && (v == ch.v0 || done < goal) && (uint32_t)n_views - v > nc - 1 - ci) {
        const uint64_t vp = (uint64_t)views[v].width * views[v].height;
        done += vp;
        ch.npix += vp;
        ch.ntiles += p->hv[v].tiles_x * ((views[v].height + 7) / 8);
        persp += views[v].projection == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
        ++v;
        ++ch.nv;
      }
      if (ci == nc - 1)
        while (v < (uint32_t)n_views) {
          const uint64_t vp = (uint64_t)views[v].width * views[v].height;
          done += vp;
          ch.npix += vp;
          ch.ntiles += p->hv[v].tiles_x * ((views[v].height + 7) / 8);
          persp += views[v].projection == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
          ++v;
          ++ch.nv;
        }
      ch.proj = persp == 0 ? VFR_TRAV_ORTHO : (persp == ch.nv ? VFR_TRAV_PERSP : VFR_TRAV_MIXED);
      p->chunks.push_back(ch);
    }
    if (hipStreamCreateWithFlags(&p->st2, hipStreamNonBlocking) != hipSuccess) {
      vfr_set_error("stream creation failed");
      rc = VFR_EHIP;
    }
    p->cev.assign(3 * p->chunks.size() + 1, nullptr);
    for (auto& e : p->cev)
      if (rc == VFR_OK && hipEventCreate(&e) != hipSuccess) { vfr_set_error("event creation failed"); rc = VFR_EHIP; }
  }
  if (rc == VFR_OK && o.schedule == VFR_SCHED_BATCHED && o.use_graph) {
    hipStream_t st = c->stream;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      vfr_set_error("graph capture failed to start");
      rc = VFR_EHIP;
    } else {
      int rc2 = enqueue_batched(p);
      if (rc2 == VFR_OK) rc2 = launch_batched_traverse(p);
      hipError_t e = hipStreamEndCapture(st, &p->graph);
      if (rc2 != VFR_OK) rc = rc2;
      else if (e != hipSuccess || hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
        vfr_set_error("graph capture failed");
        rc = VFR_EHIP;
      }
    }
  }
  if (rc != VFR_OK) { vfr_plan_destroy(p); return rc; }
  *out = p;
  return VFR_OK;
}

static double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  double pos = q * (v.size() - 1);
  size_t lo = (size_t)pos;
  size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - lo) * (v[hi] - v[lo]);
}

/* Batched execution in view chunks: chunk c's transform, raster and
 * compaction run on the context stream, its walk on the plan's second
 * stream once they are done, so the preprocessing of chunk c+1 overlaps the
 * walk of chunk c.  ms_trav sums the walk launches (HIP events on the second
 * stream). */
static int execute_pipelined(vfr_plan* p, double* ms_total, double* ms_vis, double* ms_trav) {
  hipStream_t st = p->ctx->stream, s2 = p->st2;
  const vfr_mesh* m = p->mesh;
  const size_t nc = p->chunks.size();
  uint32_t* ccount = p->counters + p->n_views + 2;
  CHECK_HIP(hipEventRecord(p->ev[0], st));
  CHECK_RC(vfr_k_clear_u32(p->owner, p->n_pix, st));
  CHECK_RC(vfr_k_clear_u32(p->counters, p->n_views + 2 + (uint32_t)nc, st));
  CHECK_RC(vfr_k_clear_u32((uint32_t*)p->stats, 2 * ST_N, st));
  for (size_t c = 0; c < nc; ++c) {
    const vfr_plan::Chunk& ch = p->chunks[c];
    CHECK_RC(vfr_k_xform_views(p->views, ch.v0, ch.nv, m->d.n_verts, (const float4*)m->d.vert, p->pverts, st));
    CHECK_RC(vfr_k_raster_ftb(p->views, ch.v0, ch.nv, m->d.n_ext, m->fvid4(), p->pverts, p->owner,
                              p->entries, p->kmax, p->n_pix, p->stats, nullptr, st));
    CHECK_RC(vfr_k_compact(p->views, p->n_views, ch.tile0, ch.ntiles, p->opts.mode, p->owner, p->owner,
                           p->entries, p->kmax, p->n_pix, p->rgba, p->steps, p->entry, p->rays + ch.pix0, ccount + c, st,
                           0, p->block_off, p->tile_mask));
    CHECK_HIP(hipEventRecord(p->cev[3 * c], st));
    CHECK_HIP(hipStreamWaitEvent(s2, p->cev[3 * c], 0));
    CHECK_HIP(hipEventRecord(p->cev[3 * c + 1], s2));
    TravArgs a = trav_args(p, p->rays + ch.pix0, ccount + c);
    a.queue = p->counters + p->n_views + 1;
    const int kern = plan_walk_kernel(p);
    if (kern != VFR_WALK_THREAD) CHECK_RC(vfr_k_clear_u32(a.queue, 1, s2));
    CHECK_RC(vfr_k_traverse(p->opts.mode, ch.proj, plan_records(p), p->opts.occupancy, a, (uint32_t)ch.npix,
                            kern, p->opts.waves_per_cu * (uint32_t)p->ctx->n_cu, s2));
    CHECK_HIP(hipEventRecord(p->cev[3 * c + 2], s2));
  }
  CHECK_HIP(hipEventRecord(p->cev[3 * nc], s2));
  CHECK_HIP(hipStreamWaitEvent(st, p->cev[3 * nc], 0));
  CHECK_HIP(hipEventRecord(p->ev[2], st));
  CHECK_HIP(hipEventSynchronize(p->ev[2]));
  float tot = 0;
  CHECK_HIP(hipEventElapsedTime(&tot, p->ev[0], p->ev[2]));
  double trav = 0;
  for (size_t c = 0; c < nc; ++c) {
    float w = 0;
    CHECK_HIP(hipEventElapsedTime(&w, p->cev[3 * c + 1], p->cev[3 * c + 2]));
    trav += w;
  }
  *ms_total = tot;
  *ms_trav = trav;
  *ms_vis = std::max(0.0, (double)tot - trav);
  return VFR_OK;
}

extern "C" int vfr_plan_execute(vfr_plan* p, vfr_stats* stats) {
  if (!p) { vfr_set_error("vfr_plan_execute: null plan"); return VFR_EINVAL; }
  if (p->mesh_version != p->mesh->version) {
    vfr_set_error("the mesh was rotated after the plan was created: create the plan again");
    return VFR_ESTATE;
  }
  vfr_ctx* c = p->ctx;
  hipStream_t st = c->stream;
  CHECK_HIP(hipSetDevice(c->device));
  double ms_total = 0, ms_vis = 0, ms_trav = 0, p50 = 0, p99 = 0;
  if (p->opts.schedule == VFR_SCHED_BATCHED && !p->chunks.empty() && (p->opts.flags & VFR_RENDER_PIPELINED)) {
    CHECK_RC(execute_pipelined(p, &ms_total, &ms_vis, &ms_trav));
  } else if (p->opts.schedule == VFR_SCHED_BATCHED) {
    CHECK_HIP(hipEventRecord(p->ev[0], st));
    if (p->gexec) {
      CHECK_HIP(hipGraphLaunch(p->gexec, st));
      CHECK_HIP(hipEventRecord(p->ev[1], st));
    } else {
      CHECK_RC(enqueue_batched(p));
      CHECK_HIP(hipEventRecord(p->ev[1], st));
      CHECK_RC(launch_batched_traverse(p));
    }
    CHECK_HIP(hipEventRecord(p->ev[2], st));
    CHECK_HIP(hipEventSynchronize(p->ev[2]));
    float a = 0, b = 0;
    CHECK_HIP(hipEventElapsedTime(&a, p->ev[0], p->ev[2]));
    ms_total = a;
    if (!p->gexec) {
      CHECK_HIP(hipEventElapsedTime(&b, p->ev[0], p->ev[1]));
      ms_vis = b;
      ms_trav = a - b;
    }
  } else {
    const vfr_mesh* m = p->mesh;
    CHECK_RC(vfr_k_clear_u32((uint32_t*)p->stats, 2 * ST_N, st));
    CHECK_RC(vfr_k_clear_u32(p->counters, p->n_views + 2, st));
    std::vector<double> frame(p->n_views), vis(p->n_views);
    CHECK_HIP(hipEventRecord(p->ev[0], st));
    for (uint32_t v = 0; v < p->n_views; ++v) {
      const DevView& hv = p->hv[v];
      uint32_t npx = hv.W * hv.H;
      uint32_t ntl = hv.tiles_x * ((hv.H + 7) / 8);
      CHECK_HIP(hipEventRecord(p->fev[2 * v], st));
      CHECK_RC(vfr_k_clear_u32(p->owner + hv.pix_off, npx, st));
      CHECK_RC(vfr_k_xform_views(p->views, v, 1, m->d.n_verts, (const float4*)m->d.vert, p->pverts, st));
      if (plan_memo_on(p))
        CHECK_RC(vfr_k_face_memo(p->views, v, 1, m->d.n_tets, m->tv4(), p->pverts, p->memo, st));
      if (p->opts.mode == VFR_MODE_REF)
        CHECK_RC(vfr_k_raster_ref(p->views, v, 1, m->d.n_ext, m->fvid4(), p->pverts, p->owner, st));
      else
        CHECK_RC(vfr_k_raster_ftb(p->views, v, 1, m->d.n_ext, m->fvid4(), p->pverts, p->owner,
                                  p->entries, p->kmax, p->n_pix, p->stats, nullptr, st));
      CHECK_RC(vfr_k_compact(p->views, p->n_views, hv.tile_off, ntl, p->opts.mode, p->owner, p->owner,
                             p->entries, p->kmax, p->n_pix, p->rgba, p->steps, p->entry, p->rays + hv.pix_off,
                             p->counters + v, st));
      CHECK_HIP(hipEventRecord(p->ev[1], st));
      TravArgs a = trav_args(p, p->rays + hv.pix_off, p->counters + v);
      a.tf_lds_base = hv.tf;
      a.tf_lds_n = 1;
      CHECK_RC(vfr_k_traverse(p->opts.mode, hv.proj == VFR_PROJ_PERSPECTIVE ? VFR_TRAV_PERSP : VFR_TRAV_ORTHO,
                              plan_records(p), 0, a, npx, 0, 0, st));
      CHECK_HIP(hipEventRecord(p->fev[2 * v + 1], st));
      CHECK_HIP(hipEventSynchronize(p->fev[2 * v + 1]));  /* a frame ends before the next starts */
      float f = 0, g = 0;
      CHECK_HIP(hipEventElapsedTime(&f, p->fev[2 * v], p->fev[2 * v + 1]));
      CHECK_HIP(hipEventElapsedTime(&g, p->fev[2 * v], p->ev[1]));
      frame[v] = f;
      vis[v] = g;
    }
    CHECK_HIP(hipEventRecord(p->ev[2], st));
    CHECK_HIP(hipEventSynchronize(p->ev[2]));
    float a = 0;
    CHECK_HIP(hipEventElapsedTime(&a, p->ev[0], p->ev[2]));
    ms_total = a;
    for (uint32_t v = 0; v < p->n_views; ++v) { ms_vis += vis[v]; ms_trav += frame[v] - vis[v]; }
    p50 = pct(frame, 0.50);
    p99 = pct(frame, 0.99);
    p->frame_ms = frame;
  }
  p->executed = true;
  /* the device counters say whether entries were dropped or a ray met a
   * non-manifold face: read them whether or not the caller wants stats */
  unsigned long long s[ST_N];
  CHECK_HIP(hipMemcpy(s, p->stats, sizeof(s), hipMemcpyDeviceToHost));
  if (stats) {
    std::vector<uint32_t> cnt(p->n_views + 2);
    CHECK_HIP(hipMemcpy(cnt.data(), p->counters, sizeof(uint32_t) * cnt.size(), hipMemcpyDeviceToHost));
    uint64_t active = 0;
    if (p->opts.schedule == VFR_SCHED_BATCHED && !p->chunks.empty() && (p->opts.flags & VFR_RENDER_PIPELINED)) {
      std::vector<uint32_t> cc(p->chunks.size());
      CHECK_HIP(hipMemcpy(cc.data(), p->counters + p->n_views + 2, sizeof(uint32_t) * cc.size(),
                          hipMemcpyDeviceToHost));
      for (uint32_t x : cc) active += x;
    } else if (p->opts.schedule == VFR_SCHED_BATCHED) {
      active = cnt[p->n_views];
    } else {
      for (uint32_t v = 0; v < p->n_views; ++v) active += cnt[v];
    }
    memset(stats, 0, sizeof(*stats));
    stats->rays = p->n_pix;
    stats->active_rays = active;
    stats->steps = s[ST_STEPS];
    stats->anomalies = s[ST_ANOM];
    stats->ms_total = ms_total;
    stats->ms_visibility = ms_vis;
    stats->ms_traversal = ms_trav;
    stats->ms_frame_p50 = p50;
    stats->ms_frame_p99 = p99;
  }
  if (s[ST_OVERFLOW]) {
    vfr_set_error("%llu pixel entries beyond %u per pixel were dropped", s[ST_OVERFLOW], p->kmax);
    return VFR_EMESH;
  }
  if (s[ST_BROKEN]) {
    vfr_set_error("%llu rays met a neighbour without a back link (non-manifold mesh)", s[ST_BROKEN]);
    return VFR_EMESH;
  }
  return VFR_OK;
}

extern "C" int vfr_plan_frame_times(const vfr_plan* p, double* out) {
  if (!p || !out) { vfr_set_error("vfr_plan_frame_times: null argument"); return VFR_EINVAL; }
  if (p->frame_ms.size() != p->n_views) { vfr_set_error("no per-view frame times (PER_VIEW plan not executed)"); return VFR_ESTATE; }
  memcpy(out, p->frame_ms.data(), sizeof(double) * p->n_views);
  return VFR_OK;
}

extern "C" int vfr_plan_read(vfr_plan* p, int view, vfr_frame_out* out) {
  if (!p || !out || view < 0 || (uint32_t)view >= p->n_views) { vfr_set_error("vfr_plan_read: bad argument"); return VFR_EINVAL; }
  if (!p->executed) { vfr_set_error("plan not executed"); return VFR_ESTATE; }
  CHECK_HIP(hipSetDevice(p->ctx->device));
  const DevView& hv = p->hv[view];
  const size_t n = (size_t)hv.W * hv.H;
  std::vector<float> rgba;
  if (out->rgba || out->rgb8) {
    float* dst = out->rgba;
    if (!dst) { rgba.resize(4 * n); dst = rgba.data(); }
    CHECK_HIP(hipMemcpy(dst, p->rgba + hv.pix_off, sizeof(float4) * n, hipMemcpyDeviceToHost));
    if (out->rgb8)
      for (size_t i = 0; i < n; ++i)
        for (int ch = 0; ch < 3; ++ch) {
          float x = dst[i * 4 + ch];
          x = x > 1.0f ? 1.0f : x;
          out->rgb8[i * 3 + ch] = (uint8_t)(int)(255.0f * x);
        }
  }
  std::vector<uint32_t> ent;
  if (out->entry_ext || out->entry_tet) {
    uint32_t* dst = out->entry_ext;
    if (!dst) { ent.resize(n); dst = ent.data(); }
    CHECK_HIP(hipMemcpy(dst, p->entry + hv.pix_off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (out->entry_tet)
      for (size_t i = 0; i < n; ++i)
        out->entry_tet[i] = dst[i] == VFR_NO_ID ? VFR_NO_ID : p->mesh->h_extf[(size_t)dst[i] * 2];
  }
  if (out->steps)
    CHECK_HIP(hipMemcpy(out->steps, p->steps + hv.pix_off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  return VFR_OK;
}

struct TuneKey {
  uint32_t mode;
  uint32_t n_tets;
  uint64_t n_pix;
  uint32_t n_views;
  uint32_t psi_table;
  bool operator==(const TuneKey& o) const {
    return mode == o.mode && n_tets == o.n_tets && n_pix == o.n_pix && n_views == o.n_views &&
           psi_table == o.psi_table;
  }
};
struct TuneVal {
  uint32_t flags, occ, wpc, refill;
};
/* The flags the autotuner chooses. */
static const uint32_t kTuned = VFR_RENDER_PERSISTENT | VFR_RENDER_WAVE_QUEUE | VFR_RENDER_PIPELINED |
                               VFR_RENDER_LDS_RECORDS | VFR_RENDER_FACE_MEMO;

/* REF plans: the whole render (the memo is filled in the preprocessing)
 * with and without the face-plane memo; the faster stays.  The memo is
 * allocated for the trial and kept only if chosen. */
static int autotune_ref(vfr_plan* p, uint32_t* chosen, double* best) {
  const uint32_t user_flags = p->opts.flags & ~kTuned;
  const bool had = p->memo != nullptr;
  if (!had && hipMalloc((void**)&p->memo, plan_memo_bytes(p)) != hipSuccess) {
    /* no room for the memo: the plain walk is the only candidate */
    (void)hipGetLastError();
    p->memo = nullptr;
  }
  vfr_stats st;
  double t[2] = {1e30, 1e30};
  int rc = VFR_OK;
  for (int m = 0; m < (p->memo ? 2 : 1) && rc == VFR_OK; ++m) {
    p->opts.flags = user_flags | (m ? VFR_RENDER_FACE_MEMO : 0u);
    for (int rep = 0; rep < 4 && rc == VFR_OK; ++rep) {
      rc = vfr_plan_execute(p, &st);
      if (rep > 0) t[m] = std::min(t[m], st.ms_total);
    }
  }
  const bool memo = rc == VFR_OK && t[1] < t[0];
  *chosen = memo ? VFR_RENDER_FACE_MEMO : 0u;
  *best = memo ? t[1] : t[0];
  p->opts.flags = user_flags | *chosen;
  if (!memo && !had && p->memo) {
    (void)hipFree(p->memo);
    p->memo = nullptr;
  }
  return rc;
}
static std::mutex g_tune_mu;
static std::vector<std::pair<TuneKey, TuneVal>> g_tune;

extern "C" int vfr_plan_autotune(vfr_plan* p, uint32_t* flags_out, uint32_t* occ_out, uint32_t* wpc_out,
                                 uint32_t* refill_out, double* best_out) {
  if (!p) { vfr_set_error("vfr_plan_autotune: null plan"); return VFR_EINVAL; }
  if (p->mesh_version != p->mesh->version) {
    vfr_set_error("the mesh was rotated after the plan was created: create the plan again");
    return VFR_ESTATE;
  }
  if (p->opts.schedule != VFR_SCHED_BATCHED) {
    vfr_set_error("autotune applies to batched plans");
    return VFR_ESTATE;
  }
  if (p->gexec) { vfr_set_error("autotune before graph capture (plan has a graph)"); return VFR_ESTATE; }
  const uint32_t table = (p->opts.flags & VFR_RENDER_PSI_TABLE) ? 1u : 0u;
  TuneKey key{p->opts.mode, p->mesh->d.n_tets, p->n_pix, p->n_views, table};
  bool cached = false;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (auto& e : g_tune)
      if (e.first == key) {
        p->opts.flags = (p->opts.flags & ~kTuned) | e.second.flags;
        p->opts.occupancy = e.second.occ;
        p->opts.waves_per_cu = e.second.wpc;
        p->opts.refill = e.second.refill;
        cached = true;
      }
  }
  double best = -1.0;
  if (cached && p->opts.mode == VFR_MODE_REF && (p->opts.flags & VFR_RENDER_FACE_MEMO) && !p->memo &&
      hipMalloc((void**)&p->memo, plan_memo_bytes(p)) != hipSuccess) {
    (void)hipGetLastError();
    p->memo = nullptr;
    p->opts.flags &= ~VFR_RENDER_FACE_MEMO;
  }
  if (!cached && p->opts.mode == VFR_MODE_REF) {
    uint32_t chosen = 0;
    const int rc = autotune_ref(p, &chosen, &best);
    if (rc != VFR_OK) return rc;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.push_back({key, TuneVal{chosen, p->opts.occupancy, p->opts.waves_per_cu, p->opts.refill}});
  } else if (!cached) {
    /* the one-ray-per-thread kernel, then persistent schedules: occupancy
     * variant x resident waves per CU (multiples of the variant's 4-SIMD
     * capacity) x refill threshold */
    static const uint32_t kOcc[] = {0, 6};
    static const uint32_t kMul[] = {1, 2};
    static const uint32_t kR[] = {16, 64};
    /* each candidate's walk is timed alone: all preprocessing first, then
     * one walk launch */
    const uint32_t user_flags = p->opts.flags & ~kTuned;
    const uint32_t base_flags = user_flags;
    best = 1e30;
    TuneVal bv{0u, p->opts.occupancy, p->opts.waves_per_cu, p->opts.refill};
    vfr_stats st;
    auto time_it = [&](double* t) -> int {
      *t = 1e30;
      for (int rep = 0; rep < 2; ++rep) {
        int rc = vfr_plan_execute(p, &st);
        if (rc != VFR_OK) return rc;
        *t = std::min(*t, st.ms_traversal);
      }
      return VFR_OK;
    };
    p->opts.flags = base_flags;
    int rc = vfr_plan_execute(p, &st); /* warm caches */
    double t = 0;
    if (rc == VFR_OK) rc = time_it(&t);
    if (rc == VFR_OK && t < best) { best = t; bv = TuneVal{0u, p->opts.occupancy, p->opts.waves_per_cu, p->opts.refill}; }
    p->opts.flags = base_flags | VFR_RENDER_WAVE_QUEUE;
    if (rc == VFR_OK) rc = time_it(&t);
    if (rc == VFR_OK && t < best) {
      best = t;
      bv = TuneVal{VFR_RENDER_WAVE_QUEUE, p->opts.occupancy, p->opts.waves_per_cu, p->opts.refill};
    }
    if (!(base_flags & VFR_RENDER_PSI_TABLE)) {
      /* the one-ray-per-thread walk with LDS-staged tet records */
      p->opts.flags = base_flags | VFR_RENDER_LDS_RECORDS;
      if (rc == VFR_OK) rc = time_it(&t);
      if (rc == VFR_OK && t < best) {
        best = t;
        bv = TuneVal{VFR_RENDER_LDS_RECORDS, p->opts.occupancy, p->opts.waves_per_cu, p->opts.refill};
      }
    }
    p->opts.flags = base_flags | VFR_RENDER_PERSISTENT;
    for (uint32_t occ : kOcc) {
      uint32_t wps = 0;
      if (rc == VFR_OK) rc = vfr_k_traverse_waves_per_simd(occ, &wps);
      for (uint32_t mul : kMul)
        for (uint32_t r : kR) {
          if (rc != VFR_OK) break;
          p->opts.occupancy = occ;
          p->opts.waves_per_cu = 4 * wps * mul;
          p->opts.refill = r;
          rc = time_it(&t);
          if (rc == VFR_OK && t < best) { best = t; bv = TuneVal{VFR_RENDER_PERSISTENT, occ, 4 * wps * mul, r}; }
        }
    }
    p->opts.flags = user_flags | bv.flags;
    p->opts.occupancy = bv.occ;
    p->opts.waves_per_cu = bv.wpc;
    p->opts.refill = bv.refill;
    if (rc == VFR_OK && !p->chunks.empty()) {
      /* end to end: one walk after all preprocessing, or view chunks */
      double serial = 1e30, piped = 1e30;
      for (int rep = 0; rc == VFR_OK && rep < 2; ++rep) {
        rc = vfr_plan_execute(p, &st);
        serial = std::min(serial, st.ms_total);
      }
      p->opts.flags |= VFR_RENDER_PIPELINED;
      for (int rep = 0; rc == VFR_OK && rep < 3; ++rep) {
        rc = vfr_plan_execute(p, &st);
        if (rep > 0) piped = std::min(piped, st.ms_total);
      }
      if (piped < serial) bv.flags |= VFR_RENDER_PIPELINED;
      p->opts.flags = user_flags | bv.flags;
    }
    if (rc != VFR_OK) return rc;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.push_back({key, bv});
  }
  if (flags_out) *flags_out = p->opts.flags & kTuned;
  if (occ_out) *occ_out = p->opts.occupancy;
  if (wpc_out) *wpc_out = p->opts.waves_per_cu;
  if (refill_out) *refill_out = p->opts.refill;
  if (best_out) *best_out = best;
  return VFR_OK;
}

extern "C" int vfr_render(vfr_ctx* c, const vfr_mesh* m, const vfr_view* views, int n_views,
                          const vfr_render_opts* opts, vfr_frame_out* outs, vfr_stats* stats) {
  vfr_plan* p = nullptr;
  CHECK_RC(vfr_plan_create(c, m, views, n_views, opts, &p));
  vfr_stats s;
  int rc = vfr_plan_execute(p, &s);
  if (rc == VFR_OK && outs)
    for (int v = 0; v < n_views && rc == VFR_OK; ++v) rc = vfr_plan_read(p, v, &outs[v]);
  if (stats) *stats = s;
  vfr_plan_destroy(p);
  return rc;
}
