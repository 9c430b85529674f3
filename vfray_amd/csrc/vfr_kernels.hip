/*
 * vfr_kernels.hip -- CDNA4 (gfx950) kernels of the VF-Ray renderer.
 *
 *   preprocessing : face keys -> radix sort -> adjacent match -> external
 *                   faces (stream compaction) -> Z-curve reorder -> packed
 *                   32-byte face records + vertex ids.  Replaces buildIncid/buildCon/
 *                   buildExtF (vfRayPreComp.cc:219-270, 424-501, 710-737).
 *   raster        : per (view, external face): visibility (the Cramer plane
 *                   test of findVisibleFaces, vfRayKernelCPU.cc:51-132), the
 *                   projected bounding box (projectVisibleFaces, :135-169) and
 *                   the per-pixel inside test of vfRay (:248-268), writing
 *                   per-pixel entries.
 *   compact       : 8x8 pixel tiles -> wave ballot -> ray list; sorts each
 *                   pixel's entry list front to back.
 *   traverse      : one ray per lane, tet walk through packed neighbour words
 *                   with the reference's exit-face test (:302-367) and
 *                   partial pre-integration (:448-482).
 *
 * All kernels are compiled with -ffp-contract=off: every float operation that
 * decides an entry face, an exit face or a step count is the same IEEE
 * operation, in the same order, as in the CPU oracle.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include <vector>

#include "../../include/vfr.h"
#include "vfr_internal.h"

#define CHECK_HIP(x)                                                          \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      vfr_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return VFR_EHIP;                                                        \
    }                                                                         \
  } while (0)

#define DEV __device__ __forceinline__

/* ------------------------------------------------------------ device maths */

/* |x| < RDELTA with x promoted to double (vfRayKernelCPU.cc:85, 206, 323).
 * (float)1e-12 is 0x2b8cbccc, just below the double 1e-12, so the double test
 * holds exactly when |x| is below the next float, 0x2b8cbccd. */
#define RDELTA_F_ABOVE 0x2b8cbccdu
DEV bool tiny(float x) { return fabsf(x) < __uint_as_float(RDELTA_F_ABOVE); }

/* (double)s <= (double)c + 1e-12 for floats s and c >= 1e-12
 * (vfRayKernelCPU.cc:342).  If s <= 2c, s - c is exact in float and the test
 * is s - c < 0x2b8cbccd; if s > 2c, the float s - c is >= c >= 0x2b8cbccd and
 * the test is false either way.  This equivalence holds while s is below 937
 * (beyond that the double sum's rounding could matter), so larger s take the
 * double path. */
DEV bool le_slack(float s, float c) {
  if (s < 937.0f) return (s - c) < __uint_as_float(RDELTA_F_ABOVE);
  return (double)s <= (double)c + 1e-12;
}

DEV uint32_t sel4(const uint4& v, uint32_t i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

DEV float4 ld_vert(const float4* vert, uint32_t i) { return vert[i]; }

/* Projective view transform; identical operation order to the oracle. */
struct ViewRegs {
  float M[12];
  uint32_t proj;
};

DEV float4 xform(const ViewRegs& v, float4 p) {
  const float* M = v.M;
  float a = ((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3];
  float b = ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7];
  float c = ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11];
  float4 r;
  if (v.proj == VFR_PROJ_ORTHO) {
    r.x = a; r.y = b; r.z = c; r.w = p.w;
  } else {
    r.z = 1.0f / c; r.x = a * r.z; r.y = b * r.z; r.w = p.w * r.z;
  }
  return r;
}

DEV ViewRegs load_view(const DevView* views, uint32_t vi) {
  ViewRegs r;
#pragma unroll
  for (int k = 0; k < 12; ++k) r.M[k] = views[vi].M[k];
  r.proj = views[vi].proj;
  return r;
}


/* P-space vertices of each view: pv[vert_off + i] = xform(view, mesh vertex i). */
/* Grid: x over the views, y over chunks of 256 vertices (looping when a mesh
 * has more chunks than the grid), so the view (and its transform) is uniform
 * across a block and read with scalar loads, and the blocks that read one
 * chunk of world vertices for every view are dispatched one after another
 * (the chunk comes from HBM once per XCD, not once per view). */
__global__ __launch_bounds__(256) void k_xform_views(const DevView* __restrict__ views, uint32_t v0,
                                                     const float4* __restrict__ wvert,
                                                     float4* __restrict__ pv) {
  const uint32_t vi = v0 + blockIdx.x;
  const uint32_t n = views[vi].n_verts;
  const ViewRegs vr = load_view(views, vi);
  const size_t out = views[vi].vert_off, in = views[vi].wvert_off;
  for (uint32_t i = blockIdx.y * blockDim.x + threadIdx.x; i < n; i += gridDim.y * blockDim.x)
    pv[out + i] = xform(vr, wvert[in + i]);
}

int vfr_k_xform_views(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_verts,
                      const float4* wvert, float4* pv, void* stream) {
  if (nv == 0 || max_verts == 0) return VFR_OK;
  const uint32_t chunks = std::min<uint32_t>((max_verts + 255) / 256, 65535u);
  hipLaunchKernelGGL(k_xform_views, dim3(nv, chunks), dim3(256), 0, (hipStream_t)stream, views, v0, wvert, pv);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* obj = (pix*2)/(real)(res-1) - 1.0 (vfRayKernelCPU.cc:250): the double
 * subtraction rounded to float equals the float subtraction, because q - 1
 * is exact in double for q in [0, 2]. */
DEV float pix_to_obj(uint32_t pix, uint32_t res) {
  float q = (float)(pix * 2u) / (float)(res - 1u);
  return q - 1.0f;
}

struct Planes {
  float z[3], s[3];
};

DEV void cramer(float c0x, float c0y, float c0z, float c1x, float c1y, float c1z, float tx,
                float ty, float tz, float det, float out[3]) {
  float tmpx = ty - tz;
  float tmpy = (ty * c1z) - (c1y * tz);
  float tmpz = (c0y * tz) - (ty * c0z);
  out[0] = ((tx * (c1y - c1z)) - (c1x * tmpx) + tmpy) / det;
  out[1] = ((c0x * tmpx) - tx * (c0y - c0z) + tmpz) / det;
  out[2] = (c0x * (-tmpy) - (c1x * tmpz) + tx * (c0y * c1z - c1y * c0z)) / det;
}

/* Face planes of vfRayKernelCPU.cc:73-122 (z and s as functions of x,y). */
DEV Planes face_planes(float4 a, float4 b, float4 c) {
  float c0x = a.x, c0y = b.x, c0z = c.x;
  float c1x = a.y, c1y = b.y, c1z = c.y;
  float det = c0x * (c1y - c1z) - c1x * (c0y - c0z) + ((c0y * c1z) - (c1y * c0z));
  det = tiny(det) ? (float)1e-12 : det;
  Planes p;
  cramer(c0x, c0y, c0z, c1x, c1y, c1z, a.z, b.z, c.z, det, p.z);
  cramer(c0x, c0y, c0z, c1x, c1y, c1z, a.w, b.w, c.w, det, p.s);
  return p;
}

DEV float plane_eval(const float p[3], float ox, float oy) { return ox * p[0] + oy * p[1] + p[2]; }

/* Face-plane memo: the z and s planes of face f of every tet (slots f, f+1,
 * f+2, the order in which the REF walk solves an exit face), per view, so a
 * walk reads the 6 coefficients instead of solving them.  Grid: x over tet
 * faces, y over views. */
__global__ __launch_bounds__(256) void k_face_memo(const DevView* __restrict__ views, uint32_t v0, uint32_t n_tets,
                                                   const uint4* __restrict__ tv, const float4* __restrict__ pverts,
                                                   float4* __restrict__ memo) {
  const uint32_t vi = v0 + blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4u * n_tets) return;
  const uint32_t t = i >> 2, f = i & 3u;
  const uint4 rv = tv[t];
  const float4* pv = pverts + views[vi].vert_off;
  const Planes pl = face_planes(pv[sel4(rv, f)], pv[sel4(rv, (f + 1u) & 3u)], pv[sel4(rv, (f + 2u) & 3u)]);
  float4* o = memo + ((size_t)vi * n_tets * 4u + i) * 2u;
  o[0] = make_float4(pl.z[0], pl.z[1], pl.z[2], pl.s[0]);
  o[1] = make_float4(pl.s[1], pl.s[2], 0.0f, 0.0f);
}

int vfr_k_face_memo(const DevView* views, uint32_t v0, uint32_t nv, uint32_t n_tets, const uint4* tv,
                    const float4* pverts, float4* memo, void* stream) {
  if (nv == 0 || n_tets == 0) return VFR_OK;
  hipLaunchKernelGGL(k_face_memo, dim3((4u * n_tets + 255u) / 256u, nv), dim3(256), 0, (hipStream_t)stream, views,
                     v0, n_tets, tv, pverts, memo);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* The memoized planes of face f of tet t in view vi. */
DEV Planes memo_planes(const float4* __restrict__ memo, uint32_t n_tets, uint32_t vi, uint32_t t, uint32_t f) {
  const float4* m = memo + (((size_t)vi * n_tets + t) * 4u + f) * 2u;
  const float4 a = m[0], b = m[1];
  Planes p;
  p.z[0] = a.x; p.z[1] = a.y; p.z[2] = a.z;
  p.s[0] = a.w; p.s[1] = b.x; p.s[2] = b.y;
  return p;
}

struct Hit {
  float cy, cz, D;
  bool sw, ok;
};

/* One candidate of the exit-face test (vfRayKernelCPU.cc:316-348). */
DEV Hit tri_test(float4 a, float4 b, float4 c, float ox, float oy) {
  Hit h;
  float v1x = b.x - a.x, v1y = b.y - a.y;
  float v2x = c.x - a.x, v2y = c.y - a.y;
  float cx = (v1x * v2y) - (v1y * v2x);
  h.ok = false;
  h.sw = false;
  h.cy = h.cz = 0.0f;
  h.D = cx;
  if (tiny(cx)) return h;
  if (cx < 0.0f) {
    float tx = v1x, ty = v1y;
    v1x = v2x; v1y = v2y; v2x = tx; v2y = ty;
    cx = -cx; h.sw = true;
  }
  float v0x = ox - a.x, v0y = oy - a.y;
  float cy = (v1x * v0y) - (v1y * v0x);
  float cz = (v0x * v2y) - (v0y * v2x);
  h.cy = cy; h.cz = cz; h.D = cx;
  h.ok = (cy >= 0.0f) && (cz >= 0.0f) && le_slack(cy + cz, cx);
  return h;
}

/* The entry inside test of vfRayKernelCPU.cc:196-268 (no slack, degenerate
 * cross replaced by RDELTA). */
DEV Hit entry_test(float4 a, float4 b, float4 c, float ox, float oy) {
  Hit h;
  float v01x = b.x - a.x, v01y = b.y - a.y, v02x = c.x - a.x, v02y = c.y - a.y;
  float cx = (v01x * v02y) - (v01y * v02x);
  cx = tiny(cx) ? (float)1e-12 : cx;
  h.sw = false;
  if (cx < 0.0f) {
    float tx = v01x, ty = v01y;
    v01x = v02x; v01y = v02y; v02x = tx; v02y = ty;
    cx = -cx; h.sw = true;
  }
  float v0px = ox - a.x, v0py = oy - a.y;
  h.cy = (v01x * v0py) - (v01y * v0px);
  h.cz = (v0px * v02y) - (v0py * v02x);
  h.D = cx;
  h.ok = (h.cy >= 0.0f) && (h.cz >= 0.0f) && ((h.cy + h.cz) <= cx);
  return h;
}

/* Value at the pixel from the inside-test crosses (weights cz/D for b and
 * cy/D for c, exchanged when the edge vectors were swapped). */
DEV float bary(float A, float B, float C, const Hit& h, float invD) {
  float wb = h.sw ? h.cy : h.cz;
  float wc = h.sw ? h.cz : h.cy;
  return A + ((wb * (B - A) + wc * (C - A)) * invD);
}

/* [0,1] clamp; NaN maps to 0 (fmaxf returns the other operand). */
DEV float clamp01(float s) { return fminf(fmaxf(s, 0.0f), 1.0f); }

DEV uint32_t fkey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
DEV float funkey(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __uint_as_float(u);
}

/* exp(x), x <= 0 (clamped at -87): Cody-Waite reduction with a two-part ln2
 * and a degree-7 Taylor polynomial on |r| <= ln2/2 in Horner form with fused
 * multiply-adds, scaled by 2^n. */
DEV float vfr_expf(float x) {
  x = fmaxf(x, -87.0f);
  const float n = __builtin_rintf(x * 1.44269504f);
  float r = __builtin_fmaf(-n, 0.693145751953125f, x);
  r = __builtin_fmaf(-n, 1.42860682030941723e-6f, r);
  float p = __builtin_fmaf(r, 0.000198412701f, 0.00138888892f);
  p = __builtin_fmaf(r, p, 0.00833333377f);
  p = __builtin_fmaf(r, p, 0.0416666679f);
  p = __builtin_fmaf(r, p, 0.166666672f);
  p = __builtin_fmaf(r, p, 0.5f);
  p = __builtin_fmaf(r, p, 1.0f);
  p = __builtin_fmaf(r, p, 1.0f);
  return __builtin_ldexpf(p, (int)n);
}

DEV uint32_t img_coord_clamped(float x, uint32_t res) {
  double f = ((double)res - 1.0) * ((double)x + 1.0) / 2.0;
  if (f <= 0.0) return 0u;
  if (f >= (double)res - 1.0) return res - 1u;
  return (uint32_t)f;
}

/* ---------------------------------------------------------- wave helpers */

DEV unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* ------------------------------------------------------------------ clear */

__global__ void k_clear(uint32_t* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}

int vfr_k_clear_u32(uint32_t* p, size_t n, void* stream) {
  if (n == 0) return VFR_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_clear, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, n);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* Plain streaming copy (16 B per lane per iteration): the reference point
 * for HBM bandwidth and for scaling the profiler's byte counters. */
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                              size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

int vfr_k_copy(const void* src, void* dst, size_t bytes, void* stream) {
  size_t n = bytes / 16;
  if (n == 0) return VFR_OK;
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, (hipStream_t)stream, (const float4*)src,
                     (float4*)dst, n);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* ----------------------------------------------------------------- raster */

/* The four P-space vertices of external face e: slots f, f+1, f+2 and the
 * one off the face, from the face's vertex ids (fvid, built with the mesh). */
DEV void face_verts(const uint4* fvid, const float4* pv, uint32_t e, float4 out[4]) {
  const uint4 q = fvid[e];
  out[0] = pv[q.x];
  out[1] = pv[q.y];
  out[2] = pv[q.z];
  out[3] = pv[q.w];
}

/* REF raster over (view, external face).  Visible faces write
 * atomicMax(owner, ext+1) into every pixel the reference's loop would write:
 * x outer, y inner over the inclusive bbox, inside test, and the column ends
 * at the first inside pixel whose entry scalar is outside [0,1]
 * (vfRayKernelCPU.cc:274).  The largest visible index wins, as the last face
 * in main.cc's loop (main.cc:209-218) overwrites the pixel. */
__global__ __launch_bounds__(256) void k_raster_ref(const DevView* __restrict__ views, uint32_t v0,
                                                    const uint4* __restrict__ fvid,
                                                    const float4* __restrict__ pv,
                                                    uint32_t* __restrict__ owner) {
  /* Column-parallel: a wave takes 64 / cpf faces, cpf lanes per face, each
   * lane the bbox columns c, c + cpf, ... of its face, every column walked
   * top-down with the reference's stop.  Columns are independent and the
   * owner is a maximum, so the pixels written are exactly those of the
   * reference's loop.  cpf (4..64) follows the view's expected face width,
   * sqrt(2 W H / faces). */
  const uint32_t vi = v0 + blockIdx.y;
  const uint32_t nE = views[vi].n_ext;
  const uint32_t W = views[vi].W, H = views[vi].H, off = views[vi].pix_off;
  const float wexp = sqrtf(2.0f * (float)W * (float)H / (float)(nE ? nE : 1u));
  uint32_t cpf = 4;
  while (cpf < 64u && (float)cpf < wexp) cpf <<= 1;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); /* known uniform */
  const uint32_t e = (blockIdx.x * (blockDim.x >> 6) + wave) * (64u / cpf) + lane / cpf;
  if (e >= nE) return;
  float4 fv[4];
  face_verts(fvid + views[vi].ext_off, pv + views[vi].vert_off, e, fv);
  Planes pl = face_planes(fv[0], fv[1], fv[2]);
  float faceZ = fv[3].x * pl.z[0] + fv[3].y * pl.z[1] + pl.z[2];
  if (faceZ < fv[3].z) return;   /* vfRayKernelCPU.cc:128 */
  if (pl.s[0] == -1.0f) return;  /* main.cc:97 drops sParams.x == -1 */
  uint32_t bx0 = W - 1, by0 = H - 1, bx1 = 0, by1 = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t ix = img_coord_clamped(fv[k].x, W), iy = img_coord_clamped(fv[k].y, H);
    bx0 = min(bx0, ix); by0 = min(by0, iy); bx1 = max(bx1, ix); by1 = max(by1, iy);
  }
  for (uint32_t px = bx0 + lane % cpf; px <= bx1; px += cpf) {
    float ox = pix_to_obj(px, W);
    for (uint32_t py = by0; py <= by1; ++py) {
      float oy = pix_to_obj(py, H);
      Hit h = entry_test(fv[0], fv[1], fv[2], ox, oy);
      if (!h.ok) continue;
      float s = plane_eval(pl.s, ox, oy);
      if ((s < 0.0f) || (s > 1.0f)) break;
      atomicMax(&owner[off + py * W + px], e + 1u);
    }
  }
}

int vfr_k_raster_ref(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_faces, const uint4* fvid,
                     const float4* pv, uint32_t* owner, void* stream) {
  if (nv == 0 || max_faces == 0) return VFR_OK;
  /* sized for one face per wave, the widest split; waves past a view's
   * faces return at once */
  hipLaunchKernelGGL(k_raster_ref, dim3((max_faces + 3) / 4, nv), dim3(256), 0, (hipStream_t)stream, views, v0,
                     fvid, pv, owner);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* FTB raster: every visible external face (front facing in P space) appends
 * the key (Z at the pixel, ext index) to each pixel its projection holds.
 * Per pixel the arithmetic is entry_test + bary's, operation for operation;
 * the order in which faces reach a pixel's list does not matter (k_compact
 * sorts each list by key).
 *
 * Work split: a wave owns 64 faces of a view, or RASTER_FPW when the view's
 * faces are large (expected area, 2 W H / faces, of RASTER_BIG pixels or
 * more) or the launch has one view, so that a single view still fills the
 * GPU.  Lanes set their
 * face up (visibility, the inside test's face part, pixel bbox).  The bbox
 * pixels of the wave's faces of up to RASTER_BAL pixels (RASTER_BAL_ONE in
 * a one-view launch) are then shared out
 * evenly, 64 per wave iteration whatever the faces' sizes; a larger face is
 * rasterized by the whole wave, 64 pixels at a time in 8x8 tiles of its
 * bbox.  Either way one atomic instruction appends to up to 64 pixel lists. */
#define RASTER_FPW 16u
#define RASTER_BIG 8u
/* bbox pixels up to which a face goes to the shared list: in a launch of
 * several views, and in a one-view launch, whose few waves cannot hide the
 * list's search (six dependent LDS reads per iteration) and whose faces of
 * more than 64 pixels keep the tile loop's lanes busy (one 1024^2 view of
 * the 64^3 mesh: 0.17 ms with 1024, 0.12 ms with 64) */
#define RASTER_BAL 1024u
#define RASTER_BAL_ONE 64u

struct RasterFace {
  uint32_t bx0, by0, bx1, by1;
  float ax, ay, v01x, v01y, v02x, v02y, cx, zA, dB, dC, invD;
  bool sw;
};

/* One pixel of a face: the inside test and, inside, the append. */
DEV void raster_pixel(const RasterFace& f, uint32_t px, uint32_t py, uint32_t W, uint32_t H, uint32_t off,
                      uint32_t fe, uint32_t* __restrict__ cnt, uint64_t* __restrict__ entries, uint32_t kmax,
                      uint64_t estride, unsigned long long& over, uint32_t* __restrict__ vflag) {
  const float ox = pix_to_obj(px, W), oy = pix_to_obj(py, H);
  const float v0px = ox - f.ax, v0py = oy - f.ay;
  const float cy = (f.v01x * v0py) - (f.v01y * v0px);
  const float cz = (v0px * f.v02y) - (v0py * f.v02x);
  if (!((cy >= 0.0f) && (cz >= 0.0f) && ((cy + cz) <= f.cx))) return;
  const float wb = f.sw ? cy : cz, wc = f.sw ? cz : cy;
  const float Zp = f.zA + ((wb * f.dB + wc * f.dC) * f.invD);
  const uint32_t gp = off + py * W + px;
  const uint32_t slot = atomicAdd(&cnt[gp], 1u);
  if (slot < kmax) {
    entries[slot * estride + gp] = ((uint64_t)fkey(Zp) << 32) | fe;
  } else {
    over++;
    if (vflag) atomicOr(vflag, VFR_VFLAG_OVERFLOW);
  }
}

__global__ __launch_bounds__(256) void k_raster_ftb(const DevView* __restrict__ views, uint32_t v0,
                                                    const uint4* __restrict__ fvid,
                                                    const float4* __restrict__ pv,
                                                    uint32_t* __restrict__ cnt,
                                                    uint64_t* __restrict__ entries, uint32_t kmax,
                                                    uint64_t estride, unsigned long long* __restrict__ stats,
                                                    uint32_t* __restrict__ vflags) {
  const uint32_t vi = v0 + blockIdx.y;
  uint32_t* const vflag = (vflags && kmax) ? vflags + vi : nullptr;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nE = views[vi].n_ext;
  const uint32_t W = views[vi].W, H = views[vi].H, off = views[vi].pix_off;
  /* a launch of one view is latency-bound: the most waves; in a batch, 64
   * faces per wave unless they are large */
  const bool spread = gridDim.y == 1 || 2ull * W * H >= (uint64_t)RASTER_BIG * nE;
  const uint32_t fpw = spread ? RASTER_FPW : 64u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); /* known uniform */
  const uint32_t e0 = (blockIdx.x * (blockDim.x >> 6) + wave) * fpw;
  if (e0 >= nE) return; /* the whole wave */
  RasterFace f{};
  bool vis = false;
  const uint32_t e = e0 + lane;
  if (lane < fpw && e < nE) {
    float4 fv[4];
    face_verts(fvid + views[vi].ext_off, pv + views[vi].vert_off, e, fv);
    const Planes pl = face_planes(fv[0], fv[1], fv[2]);
    const float faceZ = fv[3].x * pl.z[0] + fv[3].y * pl.z[1] + pl.z[2];
    double lo0 = 1e30, lo1 = 1e30, hi0 = -1e30, hi1 = -1e30;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double fx = ((double)W - 1.0) * ((double)fv[k].x + 1.0) / 2.0;
      const double fy = ((double)H - 1.0) * ((double)fv[k].y + 1.0) / 2.0;
      lo0 = fmin(lo0, fx); hi0 = fmax(hi0, fx);
      lo1 = fmin(lo1, fy); hi1 = fmax(hi1, fy);
    }
    vis = !(faceZ < fv[3].z) && !(hi0 < 0.0 || hi1 < 0.0 || lo0 > W - 1.0 || lo1 > H - 1.0);
    f.bx0 = lo0 < 0.0 ? 0u : (uint32_t)lo0;
    f.by0 = lo1 < 0.0 ? 0u : (uint32_t)lo1;
    f.bx1 = hi0 > W - 1.0 ? W - 1u : (uint32_t)hi0;
    f.by1 = hi1 > H - 1.0 ? H - 1u : (uint32_t)hi1;
    /* entry_test's face part: edge vectors from vertex 0, their cross
     * product (degenerate -> 1e-12), swapped to a positive orientation */
    f.ax = fv[0].x; f.ay = fv[0].y;
    f.v01x = fv[1].x - fv[0].x; f.v01y = fv[1].y - fv[0].y;
    f.v02x = fv[2].x - fv[0].x; f.v02y = fv[2].y - fv[0].y;
    f.cx = (f.v01x * f.v02y) - (f.v01y * f.v02x);
    f.cx = tiny(f.cx) ? (float)1e-12 : f.cx;
    f.sw = false;
    if (f.cx < 0.0f) {
      const float tx = f.v01x, ty = f.v01y;
      f.v01x = f.v02x; f.v01y = f.v02y; f.v02x = tx; f.v02y = ty;
      f.cx = -f.cx; f.sw = true;
    }
    /* bary's face part */
    f.zA = fv[0].z; f.dB = fv[1].z - fv[0].z; f.dC = fv[2].z - fv[0].z;
    f.invD = 1.0f / f.cx;
  }
  const uint32_t area = vis ? (f.bx1 - f.bx0 + 1u) * (f.by1 - f.by0 + 1u) : 0u;
  unsigned long long over = 0;
  /* Faces of up to bal pixels: the wave's (face, bbox pixel) pairs
   * form one list, face by face and row by row within a bbox, and the wave
   * takes 64 pairs at a time.  A lane finds its pair's face by binary
   * search over the faces' running pixel counts and reads the face from LDS,
   * where each lane has left its own. */
  __shared__ uint32_t s_end[4][64];
  __shared__ float4 s_face[4][64][4];
  const uint32_t bal = gridDim.y == 1 ? RASTER_BAL_ONE : RASTER_BAL;
  const uint32_t a = area <= bal ? area : 0u;
  uint32_t end = a; /* inclusive running sum over the wave's lanes */
#pragma unroll
  for (uint32_t d = 1; d < 64u; d <<= 1) {
    const uint32_t t = __shfl_up(end, d, 64);
    if (lane >= d) end += t;
  }
  const uint32_t total = __builtin_amdgcn_readlane(end, 63);
  if (total) {
    const uint32_t bw = f.bx1 - f.bx0 + 1u;
    s_end[wave][lane] = end;
    s_face[wave][lane][0] = make_float4(__uint_as_float(f.bx0 | (f.sw ? 0x80000000u : 0u)),
                                        __uint_as_float(f.by0), __uint_as_float(bw),
                                        __uint_as_float(end - a));
    s_face[wave][lane][1] = make_float4(1.0f / (float)bw, f.ax, f.ay, f.v01x);
    s_face[wave][lane][2] = make_float4(f.v01y, f.v02x, f.v02y, f.cx);
    s_face[wave][lane][3] = make_float4(f.zA, f.dB, f.dC, f.invD);
    /* the wave's own LDS rows: the wave's LDS accesses complete in order */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t base = 0; base < total; base += 64u) {
      const uint32_t item = base + lane;
      if (item >= total) break;
      /* j = number of faces whose pairs end at or before item */
      uint32_t j = 0;
#pragma unroll
      for (uint32_t st = 32; st; st >>= 1)
        if (s_end[wave][j + st - 1u] <= item) j += st;
      const float4 q0 = s_face[wave][j][0], q1 = s_face[wave][j][1];
      const float4 q2 = s_face[wave][j][2], q3 = s_face[wave][j][3];
      RasterFace g;
      const uint32_t x0 = __float_as_uint(q0.x);
      g.sw = (x0 >> 31) != 0u;
      const uint32_t gbw = __float_as_uint(q0.z);
      g.ax = q1.y; g.ay = q1.z; g.v01x = q1.w;
      g.v01y = q2.x; g.v02x = q2.y; g.v02y = q2.z; g.cx = q2.w;
      g.zA = q3.x; g.dB = q3.y; g.dC = q3.z; g.invD = q3.w;
      /* row and column of the pair in the bbox: the float quotient of two
       * integers below 2^11 is within one of the exact one */
      const uint32_t k = item - __float_as_uint(q0.w);
      uint32_t row = (uint32_t)((float)k * q1.x);
      int32_t col = (int32_t)(k - row * gbw);
      if (col < 0) { --row; col += (int32_t)gbw; }
      else if (col >= (int32_t)gbw) { ++row; col -= (int32_t)gbw; }
      raster_pixel(g, (x0 & 0x7FFFFFFFu) + (uint32_t)col, __float_as_uint(q0.y) + row, W, H, off,
                   e0 + j, cnt, entries, kmax, estride, over, vflag);
    }
  }
  /* larger faces: the whole wave, one face at a time, in 8x8 tiles */
  uint64_t todo = __ballot(vis && area > bal);
  while (todo) {
    const int j = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1ull;
    RasterFace g;
    g.bx0 = __shfl(f.bx0, j, 64); g.by0 = __shfl(f.by0, j, 64);
    g.bx1 = __shfl(f.bx1, j, 64); g.by1 = __shfl(f.by1, j, 64);
    g.ax = __shfl(f.ax, j, 64); g.ay = __shfl(f.ay, j, 64);
    g.v01x = __shfl(f.v01x, j, 64); g.v01y = __shfl(f.v01y, j, 64);
    g.v02x = __shfl(f.v02x, j, 64); g.v02y = __shfl(f.v02y, j, 64);
    g.cx = __shfl(f.cx, j, 64); g.zA = __shfl(f.zA, j, 64); g.dB = __shfl(f.dB, j, 64);
    g.dC = __shfl(f.dC, j, 64); g.invD = __shfl(f.invD, j, 64);
    g.sw = __shfl((uint32_t)f.sw, j, 64) != 0u;
    const uint32_t ge = e0 + (uint32_t)j;
    for (uint32_t ty = g.by0; ty <= g.by1; ty += 8u)
      for (uint32_t tx = g.bx0; tx <= g.bx1; tx += 8u) {
        const uint32_t px = tx + (lane & 7u), py = ty + (lane >> 3);
        if (px <= g.bx1 && py <= g.by1)
          raster_pixel(g, px, py, W, H, off, ge, cnt, entries, kmax, estride, over, vflag);
      }
  }
  if (kmax) {
    over = wave_sum_u64(over);
    if (lane == 0 && over) atomicAdd(&stats[ST_OVERFLOW], over);
  }
}

int vfr_k_raster_ftb(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_faces, const uint4* fvid,
                     const float4* pv, uint32_t* cnt, uint64_t* entries,
                     uint32_t kmax, uint64_t estride, unsigned long long* stats, uint32_t* vflags,
                     void* stream) {
  if (nv == 0 || max_faces == 0) return VFR_OK;
  const uint32_t per_block = 4u * RASTER_FPW; /* 256 threads: four waves at the smaller share */
  hipLaunchKernelGGL(k_raster_ftb, dim3((max_faces + per_block - 1) / per_block, nv), dim3(256), 0,
                     (hipStream_t)stream, views, v0, fvid, pv, cnt, entries, kmax, estride, stats, vflags);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* ---------------------------------------------------------------- compact */

#define CS(a, b)                         \
  {                                      \
    uint64_t x_ = e[a], y_ = e[b];       \
    e[a] = x_ > y_ ? x_ : y_;            \
    e[b] = x_ > y_ ? y_ : x_;            \
  }

/* TPW 8x8 tiles per wave, 16 tiles per block (one tile per wave for pass 0,
 * two for pass 1, which is limited by the loads a CU has in flight): pixels with an
 * entry become rays, inactive pixels get zero outputs, and each active
 * pixel's entry list is sorted front to back (Batcher's 19-comparator network
 * on 8 keys, descending).  Pass 0 appends the block's rays with one atomic
 * (wave ballot and a prefix over the block's waves in LDS: tile order within
 * the block, blocks in the order they finish); pass 1 only counts, keeping
 * each tile's mask and view and each block's total, for k_scan_excl and
 * k_compact_emit to write every ray at its place in tile order. */
#define COMPACT_WAVES 16u

/* Tile lt of a view's tiles_x x tiles_y 8x8 tiles, in 4x4 groups of tiles
 * (32x32 pixels, row-major groups, row-major inside a group) over the part of
 * the image the groups cover, then the tiles of the right and bottom strips
 * row-major.  The ray list follows this order, so the rays the device walks
 * at one time cover a compact patch of the image (the benchmark's 64-view
 * walk: 1.2 % faster than row-major tiles). */
DEV void tile_xy(uint32_t lt, uint32_t tiles_x, uint32_t tiles_y, uint32_t* tx, uint32_t* ty) {
  const uint32_t gx = tiles_x & ~3u, gy = tiles_y & ~3u;
  const uint32_t n_grouped = gx * gy;
  if (lt < n_grouped) {
    const uint32_t g = lt >> 4, w = lt & 15u, per_row = gx >> 2;
    *tx = (g % per_row) * 4u + (w & 3u);
    *ty = (g / per_row) * 4u + (w >> 2);
    return;
  }
  lt -= n_grouped;
  const uint32_t rw = tiles_x - gx; /* right strip: rows [0, gy) */
  if (lt < rw * gy) {
    *tx = gx + lt % rw;
    *ty = lt / rw;
    return;
  }
  lt -= rw * gy; /* bottom strip: rows [gy, tiles_y), every column */
  *tx = lt % tiles_x;
  *ty = gy + lt / tiles_x;
}
template <uint32_t TPW>
__global__ __launch_bounds__(1024) void k_compact(const DevView* __restrict__ views,
                                                  uint32_t n_views, uint32_t tile0, uint32_t ntiles,
                                                  uint32_t mode, const uint32_t* __restrict__ owner,
                                                  uint32_t* __restrict__ cnt,
                                                  uint64_t* __restrict__ entries, uint32_t kmax,
                                                  uint64_t estride,
                                                  float4* __restrict__ rgba, uint32_t* __restrict__ steps,
                                                  uint32_t* __restrict__ entry_out,
                                                  uint64_t* __restrict__ rays,
                                                  uint32_t* __restrict__ n_rays, uint32_t view_base,
                                                  uint32_t pass, uint32_t* __restrict__ block_off,
                                                  uint64_t* __restrict__ tile_mask) {
  __shared__ uint32_t wave_n[COMPACT_WAVES];
  __shared__ uint32_t block_base;
  /* the wave index through readfirstlane: the compiler then knows the tile,
   * its view and the view's fields are the same on every lane, and searches
   * and divides with scalar loads and scalar arithmetic */
  const uint32_t lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  /* lanes in Z order over the tile: lanes 4k..4k+3 are a 2x2 quad, 8k..8k+7
   * a 4x2 block (the lane groups a gather's addresses are looked up in) */
  const uint32_t lx = (lane & 1u) | ((lane >> 1) & 2u) | ((lane >> 2) & 4u);
  const uint32_t ly = ((lane >> 1) & 1u) | ((lane >> 2) & 2u) | ((lane >> 3) & 4u);
  /* the wave's TPW tiles: every pixel count is requested before any is used,
   * so that a wave has TPW loads in flight */
  uint32_t vi[TPW], pix[TPW], gp[TPW], nk[TPW];
  bool valid[TPW];
#pragma unroll
  for (uint32_t t = 0; t < TPW; ++t) {
    const uint32_t wt = blockIdx.x * COMPACT_WAVES + wave * TPW + t;
    vi[t] = 0; pix[t] = 0; gp[t] = 0; nk[t] = 0; valid[t] = false;
    if (wt < ntiles) {
      const uint32_t gt = tile0 + wt;
      /* view owning tile gt: last view with tile_off <= gt */
      uint32_t lo = 0, hi = n_views;
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (views[mid].tile_off <= gt) lo = mid; else hi = mid;
      }
      vi[t] = lo;
      const uint32_t W = views[lo].W, H = views[lo].H;
      uint32_t tx, ty;
      tile_xy(gt - views[lo].tile_off, views[lo].tiles_x, (H + 7u) / 8u, &tx, &ty);
      const uint32_t px = tx * 8u + lx, py = ty * 8u + ly;
      valid[t] = px < W && py < H;
      pix[t] = py * W + px;
      gp[t] = views[lo].pix_off + pix[t];
      if (valid[t]) nk[t] = mode == VFR_MODE_REF ? owner[gp[t]] : cnt[gp[t]];
    }
  }
  uint32_t n_act = 0;
  bool active0 = false;
  uint64_t m0 = 0;
#pragma unroll
  for (uint32_t t = 0; t < TPW; ++t) {
    const uint32_t wt = blockIdx.x * COMPACT_WAVES + wave * TPW + t;
    bool active = false;
    if (valid[t]) {
      active = nk[t] != 0u;
      if (mode != VFR_MODE_REF) {
        /* the pixel's slots, one per slot plane */
        uint64_t* ent = entries + gp[t];
        const uint32_t n = nk[t] > kmax ? kmax : nk[t];
        if (n > 1u && n <= 8u) {
          uint64_t e[8];
#pragma unroll
          for (uint32_t k = 0; k < 8; ++k) e[k] = k < n ? ent[k * estride] : 0ull;
          CS(0, 1) CS(2, 3) CS(4, 5) CS(6, 7) CS(0, 2) CS(1, 3) CS(4, 6) CS(5, 7) CS(1, 2) CS(5, 6)
          CS(0, 4) CS(1, 5) CS(2, 6) CS(3, 7) CS(2, 4) CS(3, 5) CS(1, 2) CS(3, 4) CS(5, 6)
#pragma unroll
          for (uint32_t k = 0; k < 8; ++k)
            if (k < n) ent[k * estride] = e[k];
        } else if (n > 8u) {
          /* rare long list: insertion sort in place, descending */
          for (uint32_t k = 1; k < n; ++k) {
            const uint64_t key = ent[k * estride];
            uint32_t j = k;
            while (j > 0 && ent[(j - 1) * estride] < key) { ent[j * estride] = ent[(j - 1) * estride]; --j; }
            ent[j * estride] = key;
          }
        }
      }
      if (!active) {
        rgba[gp[t]] = make_float4(0.f, 0.f, 0.f, 0.f);
        steps[gp[t]] = 0u;
        entry_out[gp[t]] = VFR_NO_ID;
      }
    }
    const uint64_t m = __ballot(active);
    /* count pass: the tile's active mask, and its view in the ntiles words
     * after the masks (the emit pass reads both instead of searching the
     * views again) */
    if (pass == 1 && lane == 0 && wt < ntiles) {
      tile_mask[wt] = m;
      reinterpret_cast<uint32_t*>(tile_mask + ntiles)[wt] = vi[t];
    }
    n_act += (uint32_t)__popcll(m);
    if (t == 0) { active0 = active; m0 = m; }
  }
  if (lane == 0) wave_n[wave] = n_act;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) tot += wave_n[w];
    if (pass == 1) block_off[blockIdx.x] = tot;
    else block_base = tot ? atomicAdd(n_rays, tot) : 0u;
  }
  __syncthreads();
  /* pass 0 (one tile per wave): the rays at the block's place in the list */
  if (TPW == 1 && active0 && pass == 0) {
    uint32_t before = block_base;
    for (uint32_t w = 0; w < wave; ++w) before += wave_n[w];
    before += (uint32_t)__popcll(m0 & ((1ull << lane) - 1ull));
    rays[before] = ((uint64_t)(vi[0] + view_base) << 32) | pix[0];
  }
}

/* Second pass of the tile-ordered compaction: two tiles per wave, as the
 * first pass, writing the rays of each tile's active mask after those of the
 * tiles before it (the block's offset from the scan, then the masks of the
 * block's earlier tiles). */
__global__ __launch_bounds__(1024) void k_compact_emit(const DevView* __restrict__ views, uint32_t n_views,
                                                       uint32_t tile0, uint32_t ntiles,
                                                       const uint64_t* __restrict__ tile_mask,
                                                       const uint32_t* __restrict__ block_off,
                                                       uint64_t* __restrict__ rays, uint32_t* __restrict__ n_rays,
                                                       uint32_t view_base) {
  /* two tiles per wave, eight waves per block: each wave's loads (the
   * masks before its tiles, then its tiles' views) are one dependent round
   * for both tiles */
  const uint32_t lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t t0 = blockIdx.x * COMPACT_WAVES, wa = t0 + 2u * wave;
  if (wa >= ntiles) return;
  const bool hb = wa + 1u < ntiles;
  uint32_t before = block_off[blockIdx.x];
  for (uint32_t u = t0; u < wa; ++u) before += (uint32_t)__popcll(tile_mask[u]);
  const uint32_t* const tview = reinterpret_cast<const uint32_t*>(tile_mask + ntiles);
  const uint64_t ma = tile_mask[wa], mb = hb ? tile_mask[wa + 1u] : 0ull;
  const uint32_t la = tview[wa], lb = hb ? tview[wa + 1u] : la;
  const uint32_t before_b = before + (uint32_t)__popcll(ma);
  if (lane == 0) {
    if (wa == ntiles - 1u) *n_rays = before_b;
    else if (wa + 1u == ntiles - 1u) *n_rays = before_b + (uint32_t)__popcll(mb);
  }
  const uint32_t Wa = views[la].W, Ha = views[la].H, Wb = views[lb].W, Hb = views[lb].H;
  const uint32_t lx = (lane & 1u) | ((lane >> 1) & 2u) | ((lane >> 2) & 4u);
  const uint32_t ly = ((lane >> 1) & 1u) | ((lane >> 2) & 2u) | ((lane >> 3) & 4u);
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t tx, ty;
  if ((ma >> lane) & 1ull) {
    tile_xy(tile0 + wa - views[la].tile_off, views[la].tiles_x, (Ha + 7u) / 8u, &tx, &ty);
    const uint32_t pix = (ty * 8u + ly) * Wa + tx * 8u + lx;
    rays[before + (uint32_t)__popcll(ma & below)] = ((uint64_t)(la + view_base) << 32) | pix;
  }
  if ((mb >> lane) & 1ull) {
    tile_xy(tile0 + wa + 1u - views[lb].tile_off, views[lb].tiles_x, (Hb + 7u) / 8u, &tx, &ty);
    const uint32_t pix = (ty * 8u + ly) * Wb + tx * 8u + lx;
    rays[before_b + (uint32_t)__popcll(mb & below)] = ((uint64_t)(lb + view_base) << 32) | pix;
  }
}

/* In-place exclusive prefix sum of v[0, n) by one block: each thread sums a
 * contiguous run, the runs are scanned in LDS, then each thread writes its
 * run's prefixes. */
__global__ __launch_bounds__(1024) void k_scan_excl(uint32_t* __restrict__ v, uint32_t n) {
  __shared__ uint32_t part[1024];
  const uint32_t per = (n + 1023u) / 1024u;
  const uint32_t b = min(n, threadIdx.x * per), e = min(n, b + per);
  uint32_t sum = 0;
  for (uint32_t i = b; i < e; ++i) sum += v[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < 1024u; d <<= 1) {
    const uint32_t t = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t c = v[i];
    v[i] = run;
    run += c;
  }
}

uint32_t vfr_compact_blocks(uint32_t ntiles) { return (ntiles + COMPACT_WAVES - 1) / COMPACT_WAVES; }

int vfr_k_compact(const DevView* views, uint32_t n_views_total, uint32_t tile0, uint32_t ntiles,
                  uint32_t mode, const uint32_t* owner, uint32_t* cnt, uint64_t* entries, uint32_t kmax,
                  uint64_t estride, float4* rgba, uint32_t* steps, uint32_t* entry_out, uint64_t* rays,
                  uint32_t* n_rays, void* stream, uint32_t view_base, uint32_t* block_off, uint64_t* tile_mask) {
  if (ntiles == 0) return VFR_OK;
  const uint32_t nb = vfr_compact_blocks(ntiles);
  hipStream_t st = (hipStream_t)stream;
  if (!block_off) {
    hipLaunchKernelGGL(k_compact<1>, dim3(nb), dim3(1024), 0, st, views, n_views_total, tile0, ntiles, mode,
                       owner, cnt, entries, kmax, estride, rgba, steps, entry_out, rays, n_rays, view_base, 0u,
                       (uint32_t*)nullptr, (uint64_t*)nullptr);
  } else {
    /* rays in tile order: count per block, scan, then write at the offsets */
    /* the count pass: two tiles per wave, eight waves per block (the
     * blocks cover the same 16 tiles as the emit pass's) */
    hipLaunchKernelGGL(k_compact<2>, dim3(nb), dim3(COMPACT_WAVES / 2 * 64), 0, st, views, n_views_total, tile0,
                       ntiles, mode, owner, cnt, entries, kmax, estride, rgba, steps, entry_out, rays, n_rays,
                       view_base, 1u, block_off, tile_mask);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, block_off, nb);
    hipLaunchKernelGGL(k_compact_emit, dim3(nb), dim3(COMPACT_WAVES / 2 * 64), 0, st, views, n_views_total, tile0,
                       ntiles, tile_mask, block_off, rays, n_rays, view_base);
  }
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

__global__ void k_max_u32(const uint32_t* __restrict__ v, size_t n, uint32_t* __restrict__ out) {
  uint32_t m = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = max(m, v[i]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor(m, d, 64));
  if ((threadIdx.x & 63u) == 0 && m) atomicMax(out, m);
}

/* Zeroes each view's pixels [pix_off, pix_off + W H) of a per-pixel array,
 * views [v0, v0 + nv), grid (max pixels / 256) x nv. */
__global__ void k_clear_view_pixels(const DevView* __restrict__ views, uint32_t v0, uint32_t* __restrict__ a) {
  const DevView& v = views[v0 + blockIdx.y];
  const uint32_t n = v.W * v.H;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[v.pix_off + i] = 0u;
}

int vfr_k_clear_view_pixels(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_pix, uint32_t* a,
                            void* stream) {
  if (nv == 0 || max_pix == 0) return VFR_OK;
  const uint32_t bx = std::min<uint32_t>((max_pix + 255u) / 256u, 1024u);
  hipLaunchKernelGGL(k_clear_view_pixels, dim3(bx, nv), dim3(256), 0, (hipStream_t)stream, views, v0, a);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

int vfr_k_max_u32(const uint32_t* cnt, size_t n, uint32_t* out, void* stream) {
  if (n == 0) return VFR_OK;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_max_u32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, cnt, n, out);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* --------------------------------------------------------------- traverse */

/* Per-lane walk state.  H[0..2] are the entry face's vertices in P space with
 * their ids; the tet record supplies the fourth. */
struct Walk {
  float4 H0, H1, H2;
  uint32_t h0, h1, h2;
  uint32_t t, fe;
};

struct Step {
  float4 A, B, C;      /* exit face vertices, the face's slot order */
  uint32_t a, b, c;    /* their ids */
  uint32_t nb;         /* packed neighbour word across the exit face */
  Hit h;
  int i;               /* 1..3, 0 = no exit found */
};

/* Component-wise selects: keeps every candidate in registers (a select of
 * whole structs lets the compiler build an indexed stack array instead). */
DEV float4 sel4f(bool c, float4 a, float4 b) {
  return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

/* REF mode's step: loads the current tet, rebuilds Q[k] = slot (fe+k)&3 from
 * the carried entry-face vertices plus the one new vertex, and runs the
 * reference's exit-face test on faces (fe+1..3)&3 in that order. */
DEV Step walk_step(const Walk& w, const TetRec* __restrict__ rec, const uint4* __restrict__ tv,
                   const float4* __restrict__ pv, float ox, float oy) {
  const uint4 rv = tv[w.t];
  const uint4 f01 = *reinterpret_cast<const uint4*>(&rec[w.t].f[0]);
  const uint4 f23 = *reinterpret_cast<const uint4*>(&rec[w.t].f[4]);
  const uint4 rn = make_uint4(f01.x, f01.z, f23.x, f23.z);
  const uint32_t fe = w.fe;
  const uint32_t q0 = sel4(rv, fe), q1 = sel4(rv, (fe + 1u) & 3u), q2 = sel4(rv, (fe + 2u) & 3u),
                 q3 = sel4(rv, (fe + 3u) & 3u);
  const float4 Q3 = pv[q3];
  const float4 Q0 = sel4f(q0 == w.h0, w.H0, sel4f(q0 == w.h1, w.H1, w.H2));
  const float4 Q1 = sel4f(q1 == w.h0, w.H0, sel4f(q1 == w.h1, w.H1, w.H2));
  const float4 Q2 = sel4f(q2 == w.h0, w.H0, sel4f(q2 == w.h1, w.H1, w.H2));
  const Hit h1 = tri_test(Q1, Q2, Q3, ox, oy);
  const Hit h2 = tri_test(Q2, Q3, Q0, ox, oy);
  const Hit h3 = tri_test(Q3, Q0, Q1, ox, oy);
  const bool c1 = h1.ok, c2 = !h1.ok && h2.ok;
  Step s;
  s.i = c1 ? 1 : (c2 ? 2 : (h3.ok ? 3 : 0));
  s.A = sel4f(c1, Q1, sel4f(c2, Q2, Q3));
  s.B = sel4f(c1, Q2, sel4f(c2, Q3, Q0));
  s.C = sel4f(c1, Q3, sel4f(c2, Q0, Q1));
  s.a = c1 ? q1 : (c2 ? q2 : q3);
  s.b = c1 ? q2 : (c2 ? q3 : q0);
  s.c = c1 ? q3 : (c2 ? q0 : q1);
  s.nb = c1 ? sel4(rn, (fe + 1u) & 3u) : (c2 ? sel4(rn, (fe + 2u) & 3u) : sel4(rn, (fe + 3u) & 3u));
  s.h.cy = c1 ? h1.cy : (c2 ? h2.cy : h3.cy);
  s.h.cz = c1 ? h1.cz : (c2 ? h2.cz : h3.cz);
  s.h.D = c1 ? h1.D : (c2 ? h2.D : h3.D);
  s.h.sw = c1 ? h1.sw : (c2 ? h2.sw : h3.sw);
  s.h.ok = s.i != 0;
  return s;
}

struct RayOut {
  float4 c;
  uint32_t steps;
  uint32_t entry;
  uint32_t anomaly;
  uint32_t broken;
};

/* Linear compositing state (vfRayKernel.cu:455-470): colour as the integer
 * 0..255 channels it accumulates, the previous sample's colour scaled to
 * 0..255 and its opacity, the accumulated opacity. */
struct LinState {
  float pr, pg, pb, pw;
  uint32_t cr, cg, cb;
  float op;
};

/* double -> uint32 as a saturating conversion: negative and NaN give 0. */
DEV uint32_t sat_u32(double v) { return v > 0.0 ? (v < 4294967296.0 ? (uint32_t)v : 0xFFFFFFFFu) : 0u; }

/* One channel of the linear integration (vfRayKernel.cu:758-772):
 * (uint)(color + (0.5 (cP + cN) transparency dz
 *                 - 0.04166666 (3 cP aP + 5 cN aP + cP aN + 3 cN aN) dz^2)),
 * with the reference's float and double sub-expressions, clamped at 255
 * (:787-792). */
DEV uint32_t lin_channel(uint32_t col, float cP, float cN, float aP, float aN, float tr, float dz) {
  const double t1 = 0.5 * (double)(cP + cN) * (double)tr * (double)dz;
  const double w = 3.0 * (double)cP * (double)aP + 5.0 * (double)cN * (double)aP + (double)(cP * aN) +
                   3.0 * (double)cN * (double)aN;
  const double t2 = 0.04166666 * w * (double)dz * (double)dz;
  const uint32_t v = sat_u32((double)col + (t1 - t2));
  return v > 255u ? 255u : v;
}

/* REF mode: the reference's per-ray loop (vfRayKernelCPU.cc:270-552) for an
 * owned pixel, with its double-promoted sub-expressions.  LIN selects the
 * linear integration of the reference's USEPARTIALPREINT 0 build
 * (vfRayKernel.cu:738-806, with the colour update the CPU version omits,
 * vfRayKernelCPU.cc:499-514); a.composite == VFR_COMPOSITE_LINEAR_OPACITY
 * adds its opacity accumulation and early termination (USEOPACITY). */
template <bool LIN>
DEV RayOut trace_ref(const TravArgs& a, uint32_t vi, uint32_t pix) {
  RayOut o;
  o.anomaly = 0; o.broken = 0;
  const float4* __restrict__ pv = a.vert + a.views[vi].vert_off;
  const TetRec* __restrict__ rec = a.rec + a.views[vi].rec_off;
  const uint4* __restrict__ tv = a.tv + a.views[vi].rec_off;
  const uint32_t* __restrict__ ext = a.ext_int + 2 * (size_t)a.views[vi].ext_off;
  const float max_edge = a.views[vi].max_edge;
  const uint32_t W = a.views[vi].W, H = a.views[vi].H, tfo = a.views[vi].tf * 256u;
  const uint32_t gp = a.views[vi].pix_off + pix;
  const uint32_t px = pix % W, py = pix / W;
  const float ox = pix_to_obj(px, W), oy = pix_to_obj(py, H);
  const uint32_t e = a.owner[gp] - 1u;
  Walk w;
  w.t = ext[e * 2];
  w.fe = ext[e * 2 + 1];
  {
    const uint4 rv = tv[w.t];
    w.h0 = sel4(rv, w.fe); w.h1 = sel4(rv, (w.fe + 1u) & 3u); w.h2 = sel4(rv, (w.fe + 2u) & 3u);
    w.H0 = pv[w.h0];
    w.H1 = pv[w.h1];
    w.H2 = pv[w.h2];
  }
  const Planes cf = face_planes(w.H0, w.H1, w.H2);
  float zPrev = plane_eval(cf.z, ox, oy);
  const float sPrev = plane_eval(cf.s, ox, oy);
  float4 cB = a.tf[tfo + min(255u, (uint32_t)(255.0f * sPrev))];
  float ax = 0.0f, ay = 0.0f, az = 0.0f, Tr = 1.0f;
  const bool opac = LIN && a.composite == VFR_COMPOSITE_LINEAR_OPACITY;
  LinState L;
  L.pr = 255.0f * cB.x; L.pg = 255.0f * cB.y; L.pb = 255.0f * cB.z; /* :463-467 */
  L.pw = opac ? cB.w : 0.0f;
  L.cr = L.cg = L.cb = 0u;
  L.op = 0.0f;
  uint32_t steps = 0;
  while (true) {
    if (++steps > VFR_MAX_STEPS) { o.anomaly = 1; break; }
    const Step s = walk_step(w, rec, tv, pv, ox, oy);
    if (s.i == 0) { o.anomaly = 1; break; }         /* :353 */
    if (s.nb == VFR_NB_BROKEN) { o.broken = 1; break; } /* :365, not emulated */
    /* exit face (fe + i) & 3: its planes from the memo, or solved here */
    const Planes nf = a.memo ? memo_planes(a.memo, a.memo_tets, vi, w.t, (w.fe + (uint32_t)s.i) & 3u)
                             : face_planes(s.A, s.B, s.C);
    const float zNext = plane_eval(nf.z, ox, oy);
    const float sNext = plane_eval(nf.s, ox, oy);
    float dz = fabsf(zPrev - zNext);
    dz = tiny(dz) ? 0.0f : dz;
    if ((sNext < 0.0f) || (sNext > 1.0f)) break;    /* :446 */
    if (LIN) {
      /* vfRayKernel.cu:738-806: dz stays in object units; without opacity
       * the transparency is 1 and the alpha weights are 0 (:755-757). */
      const float4 cN = a.tf[tfo + min(255u, (uint32_t)(255.0f * sNext))];
      const float nr = 255.0f * cN.x, ng = 255.0f * cN.y, nb = 255.0f * cN.z;
      const float nw = opac ? cN.w : 0.0f;
      const float tr = opac ? (float)(1.0 - (double)L.op) : 1.0f;
      const uint32_t r = lin_channel(L.cr, L.pr, nr, L.pw, nw, tr, dz);
      const uint32_t g = lin_channel(L.cg, L.pg, ng, L.pw, nw, tr, dz);
      const uint32_t b = lin_channel(L.cb, L.pb, nb, L.pw, nw, tr, dz);
      if (opac) L.op = (float)((double)L.op + (double)((L.pw + nw) * dz) / 2.0); /* :780 */
      L.pr = nr; L.pg = ng; L.pb = nb; L.pw = nw;
      L.cr = r; L.cg = g; L.cb = b;
      if (opac && L.op > 1.0f) { L.op = 1.0f; break; }  /* :799-803 */
      if (s.nb == VFR_NB_EXT) break;
      w.t = s.nb >> 2; w.fe = s.nb & 3u;
      w.H0 = s.A; w.H1 = s.B; w.H2 = s.C; w.h0 = s.a; w.h1 = s.b; w.h2 = s.c;
      zPrev = zNext;
      continue;
    }
    dz = dz / max_edge;
    const float4 cF = a.tf[tfo + min(255u, (uint32_t)(255.0f * sNext))];
    const float taux = cF.w * dz, tauy = cB.w * dz;
    const float zeta = (float)exp(-((double)taux * 0.5 + (double)tauy * 0.5));
    const float gx = (float)((double)taux / (1.0 + (double)taux));
    const float gy = (float)((double)tauy / (1.0 + (double)tauy));
    const float tcx = (float)((double)gx + (0.5 / 512));
    const float tcy = (float)((double)gy + (0.5 / 512));
    if ((tcx < 0.0f) || (tcx > 1.0f) || (tcy < 0.0f) || (tcy > 1.0f)) break; /* :468 */
    uint32_t ix = (uint32_t)(tcx * 512.0f), iy = (uint32_t)(tcy * 512.0f);
    ix = ix > 511u ? 511u : ix;
    iy = iy > 511u ? 511u : iy;
    const float ps = a.psi[ix * 512u + iy];
    const float d = ps - zeta;
    const float fx = (float)((double)cF.x * (1.0 - (double)ps) + (double)(cB.x * d));
    const float fy = (float)((double)cF.y * (1.0 - (double)ps) + (double)(cB.y * d));
    const float fz = (float)((double)cF.z * (1.0 - (double)ps) + (double)(cB.z * d));
    const float fw = (float)(1.0 - (double)zeta);
    ax = (float)((double)fx + (1.0 - (double)fw) * (double)ax);
    ay = (float)((double)fy + (1.0 - (double)fw) * (double)ay);
    az = (float)((double)fz + (1.0 - (double)fw) * (double)az);
    Tr *= zeta;
    cB = cF;
    if (s.nb == VFR_NB_EXT) break;                   /* :535 */
    w.t = s.nb >> 2; w.fe = s.nb & 3u;
    w.H0 = s.A; w.H1 = s.B; w.H2 = s.C; w.h0 = s.a; w.h1 = s.b; w.h2 = s.c;
    zPrev = zNext;
  }
  if (LIN) {
    /* the image byte is color * opacity (:836-838) or color (:840-842); it
     * is stored as byte / 255, which 255 * x turns back into the byte */
    const float op = opac ? L.op : 1.0f;
    const uint32_t r = opac ? (uint32_t)((float)L.cr * op) : L.cr;
    const uint32_t g = opac ? (uint32_t)((float)L.cg * op) : L.cg;
    const uint32_t b = opac ? (uint32_t)((float)L.cb * op) : L.cb;
    o.c = make_float4((float)r / 255.0f, (float)g / 255.0f, (float)b / 255.0f, op);
  } else {
    ax = ax > 1.0f ? 1.0f : ax;
    ay = ay > 1.0f ? 1.0f : ay;
    az = az > 1.0f ? 1.0f : az;
    o.c = make_float4(ax, ay, az, 1.0f - Tr);
  }
  o.steps = steps > VFR_MAX_STEPS ? VFR_MAX_STEPS : steps;
  o.entry = e;
  return o;
}

/* FTB mode as a per-lane state machine: entries front to back, segments
 * composited front to back with the front = entry point, early termination
 * at 1 - T >= term, re-entry at the first entry behind the last exit point.
 * The one-ray-per-thread kernel and the persistent work-queue kernel both
 * drive it, so they produce the same bits.
 *
 * Exit faces come from 2-D edge functions with the face vertices stored
 * relative to the pixel: for an edge u -> v the value is u.x*v.y - u.y*v.x,
 * and the same edge walked from the neighbouring tet, v -> u, gives exactly
 * the negated float, so the three faces behind an entry face cover it
 * without gaps or overlaps.  The entry face is ordered so that its edge
 * functions are >= 0 (positive orientation); the faces behind it keep that
 * orientation, so an exit face is one whose three edge functions are >= 0,
 * not all zero.  A step computes the three edges to the tet's fourth vertex;
 * the entry face's edges are carried, and the vertex the exit face leaves
 * out is replaced in place.  Depth and scalar at the exit point are the
 * edge-function-weighted average of the face's vertices (all weights >= 0).
 *
 * Two loads per step: once the exit face is known, its 8-byte link
 * (neighbour word, the neighbour's vertex off the shared face and slot map)
 * is requested from the tet's face record, and the segment is composited
 * while it travels; then the next tet's fourth vertex is requested, which the
 * next step's edge functions wait for.  The walk is bound by the vector
 * memory pipeline (one cache-line lookup per distinct line per load
 * instruction), so one 8-byte load of the exit face's link beats
 * prefetching the next tet's whole 32-byte record (two 16-byte loads) one
 * step ahead, although it puts a second load on the step's dependency chain.
 * The ray carries the record slot of each face vertex rather than its id.
 * With LDS-staged records (LDSR) the whole record is still prefetched. */
struct FtbRay {
  float4 P0, P1, P2;       /* entry face: (X - ox, Y - oy, Z, S) */
  float4 Q;                /* the tet's fourth vertex, P space (prefetched) */
  uint4 fa, fb;            /* LDSR: its face words f[0..7] (prefetched) */
  float o01, o12, o20;     /* edge functions P0->P1, P1->P2, P2->P0 (>= 0) */
  float4 cP;               /* TF colour at the entry point ...               */
  uint32_t cpb;            /* ... or its byte offset, when the TF table is in LDS */
  float ox, oy, rf, prev;  /* rf = ray length factor / max edge; prev = Z (ortho) or depth */
  float ar, ag, ab, Tr;
  uint32_t t;              /* the tet, as an index into the record arena */
  uint32_t vi, voff, gp, steps;
  uint32_t tfb;            /* byte offset of the ray's TF in the table it reads */
  uint32_t sl;             /* record slots of P0, P1, P2, Q: bytes 0..3 */
  uint32_t bits;           /* cur 0..5 | nk 6..11 | FTB_PERSP | FTB_DEAD | FTB_BROKEN */
};

#define FTB_PERSP (1u << 12)  /* perspective view                        */
#define FTB_DEAD (1u << 13)   /* ended as an anomaly (no exit face)      */
#define FTB_BROKEN (1u << 14) /* ended at a neighbour with no back link  */
DEV uint32_t ftb_cur(const FtbRay& r) { return r.bits & 63u; }
DEV uint32_t ftb_nk(const FtbRay& r) { return (r.bits >> 6) & 63u; }
DEV bool ftb_persp(const FtbRay& r) { return (r.bits & FTB_PERSP) != 0u; }

/* The slot permutations of the face records, staged in LDS once per block
 * (the caller synchronizes before the walk reads them). */
DEV void stage_slot_perms(uint32_t* sperm) {
  if (threadIdx.x < VFR_N_SLOT_PERMS) sperm[threadIdx.x] = VFR_SLOT_PERMS[threadIdx.x];
}

/* u -> v edge function of the pixel (vertices relative to the pixel). */
DEV float edge_rel(float4 u, float4 v) { return u.x * v.y - u.y * v.x; }


/* Load at a 32-bit byte offset from a kernel-argument base: one scalar base
 * register pair plus one vector offset per lane, no 64-bit address math.
 * Every array the walk reads this way stays below 4 GiB (checked when plans,
 * meshes and servers are created). */
template <class T>
DEV T ld32(const void* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

DEV float pick3(bool c1, bool c2, float a, float b, float c) { return c1 ? a : (c2 ? b : c); }
DEV uint32_t pick3u(bool c1, bool c2, uint32_t a, uint32_t b, uint32_t c) { return c1 ? a : (c2 ? b : c); }

DEV float4 rel(float4 p, float ox, float oy) { return make_float4(p.x - ox, p.y - oy, p.z, p.w); }

/* 1/x: a seed from the bit pattern (0x7EF311C3 - bits(x), within 5.1 % of
 * 1/x for every normal x of either sign) and three Newton steps
 * y += y (1 - x y) as fused multiply-adds: within 0.5005 ulp of 1/x
 * (correctly rounded for 99.98 % of inputs); NaN at +-0, which the walk
 * meets only on a step it discards (a degenerate face, D = 0).  Seven
 * full-rate instructions where IEEE division takes ten with a quarter-rate
 * reciprocal. */
DEV float vfr_rcp(float x) {
  float y = __uint_as_float(0x7EF311C3u - __float_as_uint(x));
  float e = __builtin_fmaf(-x, y, 1.0f);
  y = __builtin_fmaf(y, e, y);
  e = __builtin_fmaf(-x, y, 1.0f);
  y = __builtin_fmaf(y, e, y);
  e = __builtin_fmaf(-x, y, 1.0f);
  return __builtin_fmaf(y, e, y);
}

/* Depth value and scalar of the point with weights wA..wD (>= 0, D their
 * sum) over four vertices: ortho returns Z, perspective the depth 1/Z; both
 * from one reciprocal. */
DEV void face_values(bool persp, float D, float wA, float4 A, float wB, float4 B, float wC, float4 C,
                     float wE, float4 E, float* zd, float* s) {
  const float zn = __builtin_fmaf(wE, E.z, __builtin_fmaf(wC, C.z, __builtin_fmaf(wB, B.z, wA * A.z)));
  const float sn = __builtin_fmaf(wE, E.w, __builtin_fmaf(wC, C.w, __builtin_fmaf(wB, B.w, wA * A.w)));
  const float rc = vfr_rcp(persp ? zn : D);
  *zd = (persp ? D : zn) * rc;
  *s = clamp01(sn * rc);
}

/* exp(x) of a segment's half depth: on [-1/2, 0] (p <= 1, every segment
 * with optical depths <= 1) the degree-6 Taylor polynomial about -1/4 in
 * Horner form, within 0.81 ulp and exactly 1 at 0; below -1/2, vfr_expf. */
DEV float vfr_exp_seg(float x) {
  if (x < -0.5f) return vfr_expf(x);
  const float u = x + 0.25f;
  float e = __builtin_fmaf(u, 0.00108166772f, 0.00649000658f);
  e = __builtin_fmaf(u, e, 0.0324500315f);
  e = __builtin_fmaf(u, e, 0.129800126f);
  e = __builtin_fmaf(u, e, 0.389400393f);
  e = __builtin_fmaf(u, e, 0.778800786f);
  return __builtin_fmaf(u, e, 0.778800786f);
}

/* psi and zeta of one segment from one exponential.  psi(a, b) =
 * int_0^1 exp(-a s - (b - a) s^2 / 2) ds is the partial pre-integration
 * integral of psiGammaTable512.h:12-19 (a = optical depth at the front end, b
 * at the back) and zeta = exp(-(a + b) / 2) the segment's transparency
 * (vfRayKernelCPU.cc:458).  With p = (a + b) / 2, q = (b - a) / 2 and
 * s = 1/2 + t:
 *   psi = exp(-p/2) K(p, q),  K = int_{-1/2}^{1/2} exp(-p t + q (1/4 - t^2)) dt,
 *   zeta = exp(-p/2)^2,
 *   K = sum_{j,k} p^(2j) q^k M_jk / ((2j)! k!),  M_jk = int_{-1/2}^{1/2} t^(2j) (1/4 - t^2)^k dt,
 * kept to q-degree 5, 3, 1, 0 for p^0, p^2, p^4, p^6 (12 fused multiply-adds,
 * Horner in q, then in p^2): truncation below 4.1e-8 for a, b <= 1 (a
 * segment's depth: alpha <= 1 over at most the tet's longest edge, the
 * limits' maxEdgeLength) and 4.2e-7 at 1.5.  exp(-p/2) is vfr_exp_seg. */
DEV void psi_zeta(float a, float b, float* psi, float* zeta) {
  const float p = (a + b) * 0.5f, q = (b - a) * 0.5f;
  const float P = p * p;
  float c0 = __builtin_fmaf(q, 3.00625311e-06f, 6.61375670e-05f);
  c0 = __builtin_fmaf(q, c0, 0.00119047624f);
  c0 = __builtin_fmaf(q, c0, 0.0166666675f);
  c0 = __builtin_fmaf(q, c0, 0.166666672f);
  c0 = __builtin_fmaf(q, c0, 1.0f);
  float c1 = __builtin_fmaf(q, 1.65343918e-05f, 0.000297619059f);
  c1 = __builtin_fmaf(q, c1, 0.00416666688f);
  c1 = __builtin_fmaf(q, c1, 0.0416666679f);
  const float c2 = __builtin_fmaf(q, 3.72023824e-05f, 0.000520833360f);
  float K = __builtin_fmaf(P, 3.10019846e-06f, c2);
  K = __builtin_fmaf(P, K, c1);
  K = __builtin_fmaf(P, K, c0);
  const float E = vfr_exp_seg(p * -0.5f);
  *psi = E * K;
  *zeta = E * E;
}

/* Transfer-function colour at byte offset `b` (slot offset included): from
 * the block's LDS copy of the launch's TF slots (TFL), or from the global
 * table.  Byte offsets, so that the exit point's offset is also the next
 * step's address with no shift. */
template <int TFL>
DEV float4 tf_at(const TravArgs& a, const float4* tfl, uint32_t b) {
  return TFL ? ld32<float4>(tfl, b) : ld32<float4>(a.tf, b);
}

/* Starts the walk at entry `cur` of the pixel's sorted list; false if the
 * entry face is degenerate at the pixel.  REC: also load the tet's face
 * record into the ray (LDS-staged records). */
template <int TFL, bool REC = false>
DEV bool ftb_segment(const TravArgs& a, const float4* tfl, FtbRay& r) {
  const uint64_t key = __builtin_nontemporal_load(a.entries + ftb_cur(r) * a.estride + r.gp);
  const uint32_t e = (uint32_t)(key & 0xFFFFFFFFull) + a.views[r.vi].ext_off;
  const float4* __restrict__ pv = a.vert + r.voff;
  const uint2 tf2 = reinterpret_cast<const uint2*>(a.ext_int)[e]; /* (tet, face) */
  /* the face's vertex ids in slots fe, fe+1, fe+2 and the fourth in fe+3,
   * read beside (tet, face) rather than after it from the tet's ids */
  const uint4 fv = a.fvid[e];
  r.t = a.views[r.vi].rec_off + tf2.x;
  const uint32_t fe = tf2.y;
  if (REC) {
    r.fa = *reinterpret_cast<const uint4*>(&a.rec[r.t].f[0]);
    r.fb = *reinterpret_cast<const uint4*>(&a.rec[r.t].f[4]);
  }
  /* face fe = record slots fe, fe+1, fe+2; the fourth vertex is slot fe+3 */
  uint32_t s1 = (fe + 1u) & 3u, s2 = (fe + 2u) & 3u;
  const uint32_t s3 = (fe + 3u) & 3u;
  r.Q = pv[fv.w];
  float4 P0 = rel(pv[fv.x], r.ox, r.oy), P1 = rel(pv[fv.y], r.ox, r.oy), P2 = rel(pv[fv.z], r.ox, r.oy);
  float o01 = edge_rel(P0, P1), o12 = edge_rel(P1, P2), o20 = edge_rel(P2, P0);
  if (o01 + o12 + o20 < 0.0f) {
    /* order (P0, P2, P1): every edge function flips sign exactly */
    const float4 tP = P1; P1 = P2; P2 = tP;
    const uint32_t ts = s1; s1 = s2; s2 = ts;
    const float n01 = -o20, n12 = -o12, n20 = -o01;
    o01 = n01; o12 = n12; o20 = n20;
  }
  r.P0 = P0; r.P1 = P1; r.P2 = P2;
  r.sl = fe | (s1 << 8) | (s2 << 16) | (s3 << 24);
  r.o01 = o01; r.o12 = o12; r.o20 = o20;
  const float D = (o12 + o20) + o01;
  float sprev;
  /* P0 weighs o12, P1 o20, P2 o01 */
  face_values(ftb_persp(r), D, o12, P0, o20, P1, o01, P2, 0.0f, P2, &r.prev, &sprev);
  const uint32_t cb = r.tfb + (min(255u, (uint32_t)(255.0f * sprev)) << 4);
  if (TFL) r.cpb = cb; else r.cP = tf_at<false>(a, tfl, cb);
  return D > 0.0f;
}

/* LIN: linear compositing; the ray's colour bytes live in ar, ag, ab (exact
 * integers), its opacity in Tr, and segment lengths stay in object units. */
template <int TFL, bool LIN = false, bool REC = false>
DEV void ftb_init(const TravArgs& a, const float4* tfl, uint32_t vi, uint32_t pix, FtbRay& r) {
  const uint32_t W = a.views[vi].W, H = a.views[vi].H;
  const float k = a.views[vi].k, kx = a.views[vi].kx;
  const bool persp = a.views[vi].proj == VFR_PROJ_PERSPECTIVE;
  r.vi = vi;
  r.voff = a.views[vi].vert_off;
  /* the ray's TF as a byte offset into the table it reads: the launch's LDS
   * slots, its block's LDS slots (one per view of the block, the block's
   * first view at slot 0, which the block record after the slots names), or
   * the global table */
  if constexpr (TFL == 2) {
    r.tfb = (vi - __float_as_uint(tfl[VFR_TF_LDS_SLOTS * 256u].x)) * 4096u;
  } else {
    r.tfb = (a.views[vi].tf - (TFL ? a.tf_lds_base : 0u)) * 4096u;
  }
  r.gp = a.views[vi].pix_off + pix;
  r.ox = pix_to_obj(pix % W, W);
  r.oy = pix_to_obj(pix / W, H);
  /* perspective: world length per unit of depth along the pixel's ray,
   * |(ox tan_x, oy tan_y, 1)| with tan_x = tan(fov_y/2) W/H, tan_y = tan(fov_y/2) */
  const float ax = r.ox * kx, ay = r.oy * k;
  const float rayf = persp ? sqrtf((1.0f + ax * ax) + ay * ay) : k;
  r.rf = LIN ? rayf : rayf * a.views[vi].inv_me;
  uint32_t nk = __builtin_nontemporal_load(a.cnt + r.gp);
  nk = nk > a.kmax ? a.kmax : nk;
  r.bits = (nk << 6) | (persp ? FTB_PERSP : 0u);
  r.ar = r.ag = r.ab = 0.0f;
  r.Tr = LIN ? 0.0f : 1.0f;
  r.steps = 0;
  if (!ftb_segment<TFL, REC>(a, tfl, r)) r.bits |= FTB_DEAD;
}

/* After an exit through an external face: continue at the first entry
 * strictly behind the exit point; false when there is none (the ray is
 * finished). */
template <int TFL, bool REC = false>
DEV bool ftb_reenter(const TravArgs& a, const float4* tfl, FtbRay& r) {
  const float exitZ = ftb_persp(r) ? 1.0f / r.prev : r.prev;
  uint32_t cur = ftb_cur(r) + 1u;
  const uint32_t nk = ftb_nk(r);
  while (cur < nk && !(funkey((uint32_t)(a.entries[cur * a.estride + r.gp] >> 32)) < exitZ)) ++cur;
  if (cur >= nk) return false;
  r.bits = (r.bits & ~63u) | cur;
  if (!ftb_segment<TFL, REC>(a, tfl, r)) { r.bits |= FTB_DEAD; return false; }
  return true;
}


/* One tet of the walk on a live ray: true while the ray goes on, false
 * when it is finished.  After an exit through an external face it continues
 * at its next entry behind the exit point (ftb_reenter) within the step, so
 * the step's result is one lane mask, not a status value per lane.  `it`
 * is the step's number on the ray (1 for its first), stored in r.steps when
 * the ray ends: callers whose lanes start together pass one counter for the
 * whole wave, so no per-lane count is kept step by step.
 *
 * The body is one straight path that updates the ray state in place, so no
 * copies of the state are needed where paths join, and it is safe to run on
 * a finished ray (every index it forms stays inside the mesh arrays; the
 * result is ignored; a finished ray that tries a re-entry reads only its
 * own pixel's entry list).  The exit face's link is requested as soon as the
 * common case's exit face is known (the rare multi-candidate case requests
 * it again if it picks another face); depth, scalar, the transfer-function
 * colour, psi and the compositing run while it travels; the next tet's
 * fourth vertex is requested last, and the next step waits for it. */
/* PROJ: VFR_TRAV_*, the projection of every ray of the launch.  BIG: the
 * record arena reaches 4 GiB (VFR_BIG_RECORDS tets or more), so records are
 * addressed with 64-bit arithmetic instead of 32-bit byte offsets.  PSIT:
 * psi is looked up in the 512x512 table (nearest entry of gamma = tau /
 * (1 + tau) at both ends, the reference's convention) instead of evaluated
 * by psi_zeta. */
/* LDS-staged tet records (VFR_RENDER_LDS_RECORDS): after the slot maps
 * (sperm[0, 32)) each wave of the block owns a 64-entry table: 64 tet tags
 * and 64 owner lanes (128 words), and, further on, 64 records of 32 bytes.
 * Per step every lane whose next tet is inside the mesh tags slot (tet & 63);
 * among the lanes whose tet won the slot one becomes its owner and loads the
 * record, the others take it from LDS at the start of the next step; lanes
 * whose tet lost the slot load their own. */
#define LDSR_WORDS (32u + 4u * 128u + 4u * 512u)
DEV uint32_t* ldsr_tags(const uint32_t* sperm) {
  return const_cast<uint32_t*>(sperm) + 32 + (threadIdx.x >> 6) * 128u;
}
/* Tag accesses are relaxed wave-scope atomics: each lane reads back what the
 * whole wave wrote, so the compiler must not forward a lane's own store to
 * its load (with plain accesses it merged the stores and dropped the loads,
 * and every lane became the owner of its own record). */
DEV void tag_put(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
DEV uint32_t tag_get(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT); }
DEV uint4* ldsr_recs(const uint32_t* sperm) {
  return reinterpret_cast<uint4*>(const_cast<uint32_t*>(sperm) + 32 + 4 * 128) + (threadIdx.x >> 6) * 128u;
}
#define FTB_LSHARED (1u << 22) /* the next tet's record comes through LDS slot (bits >> 16) & 63 */
#define FTB_LOWNER (1u << 23)  /* ... and this lane loaded it */
/* The exchange of the records the wave's lanes asked for in their last step:
 * owners store what they loaded, the other lanes of the slot read it; the
 * flags are cleared. */
DEV void ldsr_exchange(const uint32_t* sperm, FtbRay& r) {
  const uint32_t h = (r.bits >> 16) & 63u;
  uint4* rr = ldsr_recs(sperm);
  if (r.bits & FTB_LOWNER) { rr[2 * h] = r.fa; rr[2 * h + 1] = r.fb; }
  if ((r.bits & (FTB_LSHARED | FTB_LOWNER)) == FTB_LSHARED) { r.fa = rr[2 * h]; r.fb = rr[2 * h + 1]; }
  r.bits &= 0xFFFFu;
}

template <int PROJ, bool BIG, int TFL, bool PSIT, bool LIN = false, bool LDSR = false>
DEV bool ftb_step(const TravArgs& a, const float4* tfl, const uint32_t* sperm, FtbRay& r, uint32_t it) {
  const bool persp = PROJ == VFR_TRAV_MIXED ? ftb_persp(r) : PROJ == VFR_TRAV_PERSP;
  /* the colour at the entry point, read first (its index is then free for
   * the exit point's) */
  const float4 cP = TFL ? tf_at<TFL>(a, tfl, r.cpb) : r.cP;
  /* Wave priority 2 from the fourth vertex's arrival to the link's request
   * and from the link's arrival to the next vertex's request, the two short
   * sequences each step's next load waits on; 0 for the rest (the
   * compositing), which can wait for issue slots. */
  __builtin_amdgcn_s_setprio(2);
  const float4 P3 = rel(r.Q, r.ox, r.oy);
  const float t0 = edge_rel(P3, r.P0), t1 = edge_rel(P3, r.P1), t2 = edge_rel(P3, r.P2);
  /* candidate faces (P3, Pi, Pj) for (i,j) = (0,1), (1,2), (2,0), edges
   * P3->Pi, Pi->Pj, Pj->P3: one is an exit face when its three edge
   * functions are >= 0 and not all zero.  o_ij >= 0 already, so the test
   * is t_i >= 0, t_j <= 0, and (rarely needed) not t_i = t_j = o_ij = 0. */
  const bool ge0 = t0 >= 0.0f, ge1 = t1 >= 0.0f, ge2 = t2 >= 0.0f;
  const bool le0 = t0 <= 0.0f, le1 = t1 <= 0.0f, le2 = t2 <= 0.0f;
  bool p1 = ge0 && le1, p2 = ge1 && le2, p3 = ge2 && le0;
  bool c1 = p1, c2 = !p1 && p2;
  /* the exit face's link: the neighbour word and the vertex word (the
   * neighbour's fourth vertex with, above bit 26, the slot map across the
   * face).  The exit face leaves out vertex k of the entry face; in the
   * record it is the face opposite k's slot s, face (s + 1) & 3.  Requested
   * at once for the face the common case picks, before the rare cases are
   * even tested, and used at the end of the step. */
  uint32_t nb, qw;
  auto link = [&](bool f1, bool f2) {
    const uint32_t sk = (r.sl >> (f1 ? 16u : (f2 ? 0u : 8u))) & 3u;
    if (LDSR) {
      const uint4 fa = r.fa, fb = r.fb;
      const bool e0 = sk == 3u, e1 = sk == 0u, e2 = sk == 1u;
      nb = pick3u(e0, e1, fa.x, fa.z, e2 ? fb.x : fb.z);
      qw = pick3u(e0, e1, fa.y, fa.w, e2 ? fb.y : fb.w);
    } else {
      const uint32_t fx = (sk + 1u) & 3u;
      const uint2 lk = BIG ? *reinterpret_cast<const uint2*>(&a.rec[r.t].f[2u * fx])
                           : ld32<uint2>(a.rec, r.t * 32u + fx * 8u);
      nb = lk.x;
      qw = lk.y;
    }
  };
  link(c1, c2);
  if ((p1 && p2) || (p2 && p3) || (p3 && p1)) {
    const bool s1 = c1, s2 = c2;
    /* rare: two edge functions are zero (a candidate is dropped when its
     * third edge is degenerate too), or the pixel is on a projected edge.
     * Two zeros always leave two candidates (t_i = t_j = 0 makes p_ij true
     * and one of the other two), so the zero rule is only needed here. */
    const bool z0 = ge0 && le0, z1 = ge1 && le1, z2 = ge2 && le2;
    p1 = p1 && !(z0 && z1 && r.o01 == 0.0f);
    p2 = p2 && !(z1 && z2 && r.o12 == 0.0f);
    p3 = p3 && !(z2 && z0 && r.o20 == 0.0f);
    c1 = p1;
    c2 = !p1 && p2;
    if ((p1 && p2) || (p2 && p3) || (p3 && p1)) {
      /* exit at the point farthest along the ray (smallest Z), ties in
       * candidate order */
      const float inf = __uint_as_float(0x7f800000u);
      const float z1 = p1 ? P3.z + (-t1 * (r.P0.z - P3.z) + t0 * (r.P1.z - P3.z)) / (t0 + r.o01 - t1) : inf;
      const float z2 = p2 ? P3.z + (-t2 * (r.P1.z - P3.z) + t1 * (r.P2.z - P3.z)) / (t1 + r.o12 - t2) : inf;
      const float z3 = p3 ? P3.z + (-t0 * (r.P2.z - P3.z) + t2 * (r.P0.z - P3.z)) / (t2 + r.o20 - t0) : inf;
      c1 = p1 && z1 <= z2 && z1 <= z3;
      c2 = !c1 && p2 && z2 <= z3;
    }
    if (c1 != s1 || c2 != s2) link(c1, c2); /* the face the rare case picked */
  }
  /* no exit face or the step bound: the step is not composited */
  const bool skip = it > VFR_MAX_STEPS || !(p1 || p2 || p3);
  const bool k0 = c2, k1 = !c1 && !c2, k2 = c1;
  __builtin_amdgcn_s_setprio(0);
  /* exit face (P3, Pi, Pj): P3 weighs o_ij, Pi weighs -t_j, Pj weighs t_i,
   * the left-out vertex Pk weighs 0 */
  const float w3 = pick3(c1, c2, r.o01, r.o12, r.o20);
  const float w0 = c1 ? -t1 : (c2 ? 0.0f : t2);
  const float w1 = c1 ? t0 : (c2 ? -t2 : 0.0f);
  const float w2 = c1 ? 0.0f : (c2 ? t1 : -t0);
  const float D = ((w3 + w0) + w1) + w2;
  float zd, snext;
  face_values(persp, D, w3, P3, w0, r.P0, w1, r.P1, w2, r.P2, &zd, &snext);
  /* the colour at the exit point; psi is evaluated in registers (with PSIT
   * it is a table entry, a second load waited for within the step).  A
   * skipped step composites nothing: with dl = 0 both depths are 0, psi =
   * zeta = 1 and the update leaves colour and transparency unchanged (in
   * linear compositing, both terms vanish). */
  const uint32_t cb = r.tfb + ((uint32_t)(255.0f * snext) << 4); /* snext is in [0, 1] */
  const float4 cN = tf_at<TFL>(a, tfl, cb);
  /* (the oracle's max(.., 0) changes nothing here: a step with an exit face
   * has D > 0 and finite depths, and a skipped one is zeroed) */
  const float dl = skip ? 0.0f : fabsf(zd - r.prev) * r.rf;
  if (LIN) {
    /* the reference's linear integration (vfRayKernel.cu:738-806, as in
     * trace_ref) over a segment of dl object units, front to back */
    const bool opac = a.composite == VFR_COMPOSITE_LINEAR_OPACITY;
    const float pw = opac ? cP.w : 0.0f, nw = opac ? cN.w : 0.0f;
    const float tr = opac ? (float)(1.0 - (double)r.Tr) : 1.0f;
    r.ar = (float)lin_channel((uint32_t)r.ar, 255.0f * cP.x, 255.0f * cN.x, pw, nw, tr, dl);
    r.ag = (float)lin_channel((uint32_t)r.ag, 255.0f * cP.y, 255.0f * cN.y, pw, nw, tr, dl);
    r.ab = (float)lin_channel((uint32_t)r.ab, 255.0f * cP.z, 255.0f * cN.z, pw, nw, tr, dl);
    if (opac) r.Tr = (float)((double)r.Tr + (double)((pw + nw) * dl) / 2.0);
  }
  float ps = 1.0f, zeta = 1.0f;
  if (!LIN) {
    const float tf_ = cP.w * dl, tb_ = cN.w * dl;
    if (PSIT) {
      /* 512 * gamma = 512 * tau / (1 + tau) for both ends with one division */
      const float uf = 1.0f + tf_, ub = 1.0f + tb_;
      const float q = 512.0f / (uf * ub);
      const float gf = (tf_ * ub) * q, gb = (tb_ * uf) * q;
      ps = a.psi[min(511u, (uint32_t)(gf + 0.5f)) * 512u + min(511u, (uint32_t)(gb + 0.5f))];
      __builtin_amdgcn_sched_barrier(0);
      zeta = vfr_expf(-(tf_ + tb_) * 0.5f);
    } else {
      psi_zeta(tf_, tb_, &ps, &zeta);
    }
  }
  /* the exit face becomes the entry face: vertex k is replaced by P3 */
  const float n01 = pick3(k0, k1, t1, -t0, r.o01);
  const float n12 = pick3(k1, k2, t2, -t1, r.o12);
  const float n20 = pick3(k2, k0, t0, -t2, r.o20);
  r.P0 = sel4f(k0, P3, r.P0);
  r.P1 = sel4f(k1, P3, r.P1);
  r.P2 = sel4f(k2, P3, r.P2);
  r.o01 = n01; r.o12 = n12; r.o20 = n20;
  /* composite the segment */
  if (!LIN) {
    const float w1c = 1.0f - ps, w2c = ps - zeta;
    const float ar = __builtin_fmaf(r.Tr, __builtin_fmaf(cP.x, w1c, cN.x * w2c), r.ar);
    const float ag = __builtin_fmaf(r.Tr, __builtin_fmaf(cP.y, w1c, cN.y * w2c), r.ag);
    const float ab = __builtin_fmaf(r.Tr, __builtin_fmaf(cP.z, w1c, cN.z * w2c), r.ab);
    r.ar = ar;
    r.ag = ag;
    r.ab = ab;
    r.Tr = r.Tr * zeta;
  }
  if (TFL) r.cpb = cb; else r.cP = cN;
  r.prev = zd;
  /* Empty statements that take the composited colour and transparency and
   * the new entry face as register inputs: they make the compiler finish
   * both here, before the fourth vertex is requested.  Left to itself it
   * moved the compositing past the step's branch into the loop's back edge
   * and updated the entry face after the request, so the old transparency
   * and the old fourth vertex were still live there: it copied the loaded
   * vertex into the loop's registers (waiting for the load at once) and
   * copied the transparency every step.  Without the copies the step is
   * about 2.5 % faster, with the same results. */
  if (!LIN) asm volatile("" ::"v"(r.ar), "v"(r.ag), "v"(r.ab), "v"(r.Tr));
  asm volatile("" ::"v"(r.P0.x), "v"(r.P0.y), "v"(r.P0.z), "v"(r.P0.w), "v"(r.P1.x), "v"(r.P1.y),
               "v"(r.P1.z), "v"(r.P1.w), "v"(r.P2.x), "v"(r.P2.y), "v"(r.P2.z), "v"(r.P2.w));
  /* the link: the next tet (the current one again at a boundary, unused),
   * its fourth vertex, requested here and waited for by the next step, and
   * the slots in its record: vertex k and the fourth vertex trade places (P3
   * replaces vertex k, the neighbour's vertex off the face becomes the
   * fourth, and the map sends k's slot to that vertex's slot), then every
   * byte goes through the map.  Neighbour words index the record arena (a
   * server rebases each resident mesh's words when it copies them in). */
  __builtin_amdgcn_s_setprio(2);
  const uint32_t i3n = qw & VFR_REC_VERT_MASK;
  const uint32_t smap = sperm[qw >> VFR_REC_VERT_BITS];
  const uint32_t tn = nb < VFR_NB_BROKEN ? (nb >> 2) : r.t;
  r.Q = ld32<float4>(a.vert, (r.voff + i3n) * 16u);
  __builtin_amdgcn_s_setprio(0);
  const uint32_t swp = c1 ? 0x02030100u : (c2 ? 0x00020103u : 0x01020300u);
  r.sl = __builtin_amdgcn_perm(smap, smap, __builtin_amdgcn_perm(r.sl, r.sl, swp));
  if (LDSR) {
    /* the next tet's record; lanes at a boundary or on a skipped step keep to
     * themselves */
    bool own = true;
    if (nb < VFR_NB_BROKEN && !skip) {
      const uint32_t lane = threadIdx.x & 63u, h = tn & 63u;
      uint32_t* tag = ldsr_tags(sperm);
      /* the tet that wins slot h, then one lane among those asking for it */
      tag_put(tag + h, tn);
      if (tag_get(tag + h) == tn) {
        tag_put(tag + 64 + h, lane);
        own = tag_get(tag + 64 + h) == lane;
        r.bits |= (h << 16) | FTB_LSHARED | (own ? FTB_LOWNER : 0u);
      }
    }
    if (own) {
      r.fa = ld32<uint4>(a.rec, tn * 32u);
      r.fb = ld32<uint4>(a.rec, tn * 32u + 16u);
    }
  }
  r.t = tn;
  /* linear: the reference's stop once the opacity passes 1 (vfRayKernel.cu:
   * 799-803), opacity clamped to 1 */
  const bool lstop = LIN && a.composite == VFR_COMPOSITE_LINEAR_OPACITY && r.Tr > 1.0f;
  if (lstop) r.Tr = 1.0f;
  /* 1 - Tr >= term, as one compare: 1 - T rounded to float never grows with
   * T, so it holds exactly for T <= term_tr */
  const bool term = LIN ? lstop : r.Tr <= a.term_tr;
  if (skip || term || nb >= VFR_NB_BROKEN) {
    /* rare: the ray ends or leaves through the boundary */
    r.steps = it;
    if (skip) { r.bits |= FTB_DEAD; return false; }
    if (term) return false;
    if (nb == VFR_NB_BROKEN) { r.bits |= FTB_BROKEN; return false; }
    return ftb_reenter<TFL, LDSR>(a, tfl, r);
  }
  return true;
}

template <bool LIN = false>
DEV RayOut ftb_out(const TravArgs& a, const FtbRay& r) {
  RayOut o;
  if (LIN) {
    /* the image byte is color * opacity with opacity, color without
     * (vfRayKernel.cu:836-842), stored as byte / 255 */
    const bool opac = a.composite == VFR_COMPOSITE_LINEAR_OPACITY;
    const float op = opac ? r.Tr : 1.0f;
    const uint32_t cr = (uint32_t)r.ar, cg = (uint32_t)r.ag, cb = (uint32_t)r.ab;
    const uint32_t br = opac ? (uint32_t)((float)cr * op) : cr;
    const uint32_t bg = opac ? (uint32_t)((float)cg * op) : cg;
    const uint32_t bb = opac ? (uint32_t)((float)cb * op) : cb;
    o.c = make_float4((float)br / 255.0f, (float)bg / 255.0f, (float)bb / 255.0f, op);
  } else {
    o.c = make_float4(r.ar, r.ag, r.ab, 1.0f - r.Tr);
  }
  o.steps = r.steps > VFR_MAX_STEPS ? VFR_MAX_STEPS : r.steps;
  o.entry = (uint32_t)(a.entries[r.gp] & 0xFFFFFFFFull);
  o.anomaly = (r.bits & FTB_DEAD) ? 1u : 0u;
  o.broken = (r.bits & FTB_BROKEN) ? 1u : 0u;
  return o;
}

template <int PROJ, bool BIG, int TFL, bool PSIT, bool LIN = false, bool LDSR = false>
DEV RayOut trace_ftb(const TravArgs& a, const float4* tfl, const uint32_t* sperm, uint32_t vi, uint32_t pix) {
  FtbRay r;
  ftb_init<TFL, LIN, LDSR>(a, tfl, vi, pix, r);
  if (LDSR) {
    /* a lane whose ray has just ended takes part in one more exchange, so
     * that a record it loaded for other lanes reaches them */
    bool live = !(r.bits & FTB_DEAD);
    uint32_t it = 0;
    while (true) {
      ldsr_exchange(sperm, r);
      if (!live) break;
      live = ftb_step<PROJ, BIG, TFL, PSIT, LIN, true>(a, tfl, sperm, r, ++it);
    }
  } else if (!(r.bits & FTB_DEAD)) {
    uint32_t it = 0;
    while (ftb_step<PROJ, BIG, TFL, PSIT, LIN>(a, tfl, sperm, r, ++it)) {
    }
  }
  return ftb_out<LIN>(a, r);
}

template <uint32_t MODE, int PROJ, bool BIG, int TFL, bool PSIT, bool LDSR = false>
DEV RayOut trace(const TravArgs& a, const float4* tfl, const uint32_t* sperm, uint32_t vi, uint32_t pix) {
  if constexpr (MODE == VFR_MODE_REF) return trace_ref<false>(a, vi, pix);
  else if constexpr (MODE == VFR_MODE_REF_LINEAR) return trace_ref<true>(a, vi, pix);
  else if constexpr (MODE == VFR_MODE_FTB_LINEAR) return trace_ftb<PROJ, BIG, TFL, false, true>(a, tfl, sperm, vi, pix);
  else return trace_ftb<PROJ, BIG, TFL, PSIT, false, LDSR>(a, tfl, sperm, vi, pix);
}

/* Per-ray inputs and outputs are touched once: non-temporal accesses keep
 * them from evicting the tet records and vertices from L2 and the
 * Infinity Cache. */
typedef float vfr_f4v __attribute__((ext_vector_type(4)));
DEV void store_ray(const TravArgs& a, uint32_t gp, const RayOut& o) {
  const vfr_f4v c = {o.c.x, o.c.y, o.c.z, o.c.w};
  __builtin_nontemporal_store(c, reinterpret_cast<vfr_f4v*>(a.rgba) + gp);
  __builtin_nontemporal_store(o.steps, a.steps + gp);
  __builtin_nontemporal_store(o.entry, a.entry_out + gp);
}

/* One ray per thread over the compacted ray list. */
/* TFL: the launch's transfer-function slots [tf_lds_base, +tf_lds_n) are
 * staged in LDS once per block (every view of the launch uses one of them). */
/* LDSR: the FTB walk with LDS-staged tet records (256-thread blocks).
 * Waves per SIMD: 7 for the FTB walk (72 VGPRs, no spills; 6 measured 1.6 %
 * slower and 8 spills), 6 with LDS-staged records, 5 for the REF walk. */
#define VFR_TRAV_WAVES(MODE, LDSR) ((MODE) == VFR_MODE_FTB ? ((LDSR) ? 6 : 7) : 5)
template <uint32_t MODE, int PROJ, bool BIG, int TFL, bool PSIT, bool LDSR = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VFR_TRAV_WAVES(MODE, LDSR),
                                                                     VFR_TRAV_WAVES(MODE, LDSR)))) void k_trav(
    TravArgs a) {
  /* TF slots, then (per-block staging) the block record: its first view */
  __shared__ float4 tfs[TFL == 2 ? VFR_TF_LDS_SLOTS * 256 + 1 : (TFL ? VFR_TF_LDS_SLOTS * 256 : 1)];
  __shared__ __align__(16) uint32_t sperm[LDSR ? LDSR_WORDS : VFR_N_SLOT_PERMS];
  constexpr bool ftb = MODE == VFR_MODE_FTB || MODE == VFR_MODE_FTB_LINEAR;
  /* Blocks go to the 8 XCDs round robin (block b to XCD b % 8).  Within each
   * run of 128 blocks, XCD x walks the consecutive ray blocks 16x .. 16x+15
   * (neighbouring screen tiles), so its L2 holds one region of the mesh
   * instead of 16 scattered ones (FTB walk 1-2 % faster; one contiguous
   * eighth of the ray list per XCD was slower: the XCDs finish unevenly). */
  uint32_t lb = blockIdx.x;
  if (ftb && lb < (gridDim.x & ~127u)) {
    const uint32_t w = lb & 127u;
    lb = (lb & ~127u) | ((w & 7u) << 4) | (w >> 3);
  }
  const uint32_t i = lb * blockDim.x + threadIdx.x;
  const uint32_t n = *a.n_rays;
  const uint64_t ray = i < n ? __builtin_nontemporal_load(a.rays + i) : 0ull;
  const uint32_t vi = (uint32_t)(ray >> 32), pix = (uint32_t)ray;
  if (ftb) stage_slot_perms(sperm);
  bool bglob = false;
  if constexpr (TFL == 2) {
    /* the TFs of the views the block's rays belong to, one LDS slot per view
     * from the least to the greatest (the list is grouped by view, so a
     * block spans one or two views unless views are tiny), when they fit;
     * else the block reads the global table */
    __shared__ uint32_t vlo[4], vhi[4], vused;
    uint32_t lo = i < n ? vi : 0xFFFFFFFFu, hi = i < n ? vi : 0u;
    for (int d = 32; d >= 1; d >>= 1) {
      lo = min(lo, (uint32_t)__shfl_xor(lo, d, 64));
      hi = max(hi, (uint32_t)__shfl_xor(hi, d, 64));
    }
    if ((threadIdx.x & 63u) == 0) { vlo[threadIdx.x >> 6] = lo; vhi[threadIdx.x >> 6] = hi; }
    if (threadIdx.x == 0) vused = 0u;
    __syncthreads();
    lo = min(min(vlo[0], vlo[1]), min(vlo[2], vlo[3]));
    hi = max(max(vhi[0], vhi[1]), max(vhi[2], vhi[3]));
    const uint32_t nv = lo <= hi ? hi - lo + 1u : 0u;
    const bool glob = nv > VFR_TF_LDS_SLOTS;
    bglob = __builtin_amdgcn_readfirstlane(glob ? 1 : 0) != 0; /* the same on every lane */
    /* only the views some ray of the block has are read (view slots between
     * them may be unused, as in a server's shared view arena) */
    if (!glob && i < n) atomicOr(&vused, 1u << (vi - lo));
    __syncthreads();
    const uint32_t used = vused;
    if (!glob)
      for (uint32_t k = threadIdx.x; k < nv * 256u; k += blockDim.x)
        if (used & (1u << (k >> 8))) tfs[k] = a.tf[a.views[lo + (k >> 8)].tf * 256u + (k & 255u)];
    if (threadIdx.x == 0)
      tfs[VFR_TF_LDS_SLOTS * 256] = make_float4(__uint_as_float(lo), 0.f, 0.f, 0.f);
  } else if constexpr (TFL == 1) {
    for (uint32_t k = threadIdx.x; k < a.tf_lds_n * 256u; k += blockDim.x) tfs[k] = a.tf[a.tf_lds_base * 256u + k];
  }
  if (TFL || ftb) __syncthreads();
  unsigned long long st = 0, an = 0, br = 0;
  if (i < n) {
    /* per-block staging: the block's choice picks one of two walks once, so
     * no step asks which table it reads */
    RayOut o;
    if constexpr (TFL == 2)
      o = bglob ? trace<MODE, PROJ, BIG, 0, PSIT, LDSR>(a, tfs, sperm, vi, pix)
                : trace<MODE, PROJ, BIG, 2, PSIT, LDSR>(a, tfs, sperm, vi, pix);
    else
      o = trace<MODE, PROJ, BIG, TFL, PSIT, LDSR>(a, tfs, sperm, vi, pix);
    store_ray(a, a.views[vi].pix_off + pix, o);
    st = o.steps; an = o.anomaly; br = o.broken;
    if (o.broken && a.vflags) atomicOr(a.vflags + vi, VFR_VFLAG_BROKEN);
  }
  st = wave_sum_u64(st);
  an = wave_sum_u64(an);
  br = wave_sum_u64(br);
  if ((threadIdx.x & 63u) == 0) {
    if (st) atomicAdd(&a.stats[ST_STEPS], st);
    if (an) atomicAdd(&a.stats[ST_ANOM], an);
    if (br) atomicAdd(&a.stats[ST_BROKEN], br);
  }
}

/* Wave work-queue traversal: a fixed grid of waves, each of which takes the
 * next 64 rays of the list with one atomic (one ray per lane, as k_trav) and
 * walks them to the end, then takes the next 64, until the list is drained.
 * A wave whose rays end goes on at once, where a k_trav block keeps its
 * dispatch slot until its slowest wave ends; the transfer function is
 * staged in LDS once per block.  Every wave reaches the drained queue, so
 * the grid ends. */
template <int PROJ, bool BIG, int TFL, bool PSIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_trav_wq(TravArgs a) {
  __shared__ float4 tfs[TFL ? VFR_TF_LDS_SLOTS * 256 : 1];
  __shared__ uint32_t sperm[VFR_N_SLOT_PERMS];
  stage_slot_perms(sperm);
  if (TFL) {
    for (uint32_t k = threadIdx.x; k < a.tf_lds_n * 256u; k += blockDim.x) tfs[k] = a.tf[a.tf_lds_base * 256u + k];
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = *a.n_rays;
  unsigned long long st = 0, an = 0, br = 0;
  while (true) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(a.queue, 64u);
    base = __shfl(base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane;
    if (i < n) {
      const uint64_t ray = a.rays[i];
      const uint32_t vi = (uint32_t)(ray >> 32), pix = (uint32_t)ray;
      const RayOut o = trace<VFR_MODE_FTB, PROJ, BIG, TFL, PSIT>(a, tfs, sperm, vi, pix);
      store_ray(a, a.views[vi].pix_off + pix, o);
      st += o.steps; an += o.anomaly; br += o.broken;
      if (o.broken && a.vflags) atomicOr(a.vflags + vi, VFR_VFLAG_BROKEN);
    }
  }
  st = wave_sum_u64(st);
  an = wave_sum_u64(an);
  br = wave_sum_u64(br);
  if (lane == 0) {
    if (st) atomicAdd(&a.stats[ST_STEPS], st);
    if (an) atomicAdd(&a.stats[ST_ANOM], an);
    if (br) atomicAdd(&a.stats[ST_BROKEN], br);
  }
}

/* Persistent work-queue traversal: a fixed grid of waves pulls rays from the
 * compacted list.  Whenever at least `refill` lanes of a wave are idle (their
 * rays ended or left through a concave boundary for good) the wave claims
 * that many rays with one atomic and hands them to the idle lanes (ballot +
 * prefix count), then advances every lane `burst` tets.  A wave exits once
 * the queue is drained and all its lanes are idle; every ray ends within
 * VFR_MAX_STEPS tets, so every wave reaches that exit. */
template <int WPE, bool PSIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_trav_persist(
    TravArgs a, uint32_t burst) {
  __shared__ uint32_t sperm[VFR_N_SLOT_PERMS];
  stage_slot_perms(sperm);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = *a.n_rays;
  const uint32_t refill = a.refill < 1u ? 1u : (a.refill > 64u ? 64u : a.refill);
  if (n == 0) return;
  /* every lane starts on a copy of the first ray, so that an idle lane's
   * state is a valid one (its steps are never stored) */
  FtbRay r;
  {
    const uint64_t ray = a.rays[0];
    ftb_init<false>(a, nullptr, (uint32_t)(ray >> 32), (uint32_t)ray, r);
  }
  bool active = false, drained = false;
  uint32_t st = 0, an = 0, br = 0;
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t ni = (uint32_t)__popcll(idle);
    if (!drained && ni >= refill) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(a.queue, ni);
      base = __shfl(base, 0, 64);
      if (base >= n) {
        drained = true;
      } else if (!active) {
        const uint32_t idx = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        if (idx < n) {
          const uint64_t ray = a.rays[idx];
          ftb_init<false>(a, nullptr, (uint32_t)(ray >> 32), (uint32_t)ray, r);
          active = true;
          if (r.bits & FTB_DEAD) {
            /* degenerate entry face: the ray ends before its first step */
            const RayOut o = ftb_out(a, r);
            store_ray(a, r.gp, o);
            an += o.anomaly;
            active = false;
          }
        }
      }
    }
    if (__ballot(active) == 0ull) {
      if (drained) break;
      continue;
    }
    /* every lane steps (an idle lane's stale ray stays inside the mesh
     * arrays and its result is ignored), so the walk is one straight path */
    for (uint32_t k = 0; k < burst; ++k) {
      /* an idle lane's stale ray may also try a re-entry: it reads only its
       * own pixel's entry list, and its result is ignored */
      const bool go = ftb_step<VFR_TRAV_MIXED, true, false, PSIT>(a, nullptr, sperm, r, ++r.steps);
      if (active && !go) {
        const RayOut o = ftb_out(a, r);
        store_ray(a, r.gp, o);
        st += o.steps; an += o.anomaly; br += o.broken;
        if (o.broken && a.vflags) atomicOr(a.vflags + r.vi, VFR_VFLAG_BROKEN);
        active = false;
      }
    }
  }
  const unsigned long long ws = wave_sum_u64(st), wa = wave_sum_u64(an), wb = wave_sum_u64(br);
  if (lane == 0) {
    if (ws) atomicAdd(&a.stats[ST_STEPS], ws);
    if (wa) atomicAdd(&a.stats[ST_ANOM], wa);
    if (wb) atomicAdd(&a.stats[ST_BROKEN], wb);
  }
}

template <int WPE>
static void launch_persist(uint32_t blocks, void* stream, const TravArgs& a, uint32_t burst) {
  if (a.psi_table)
    hipLaunchKernelGGL((k_trav_persist<WPE, true>), dim3(blocks), dim3(256), 0, (hipStream_t)stream, a, burst);
  else
    hipLaunchKernelGGL((k_trav_persist<WPE, false>), dim3(blocks), dim3(256), 0, (hipStream_t)stream, a, burst);
}

int vfr_k_traverse_waves_per_simd(uint32_t occupancy, uint32_t* out) {
  int blocks = 0;
  hipError_t e;
  if (occupancy >= 6) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trav_persist<6, false>, 256, 0);
  else if (occupancy == 5) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trav_persist<5, false>, 256, 0);
  else if (occupancy == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trav_persist<4, false>, 256, 0);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trav_persist<1, false>, 256, 0);
  CHECK_HIP(e);
  /* a 256-thread block is 4 waves, one per SIMD of the CU */
  *out = blocks < 1 ? 1u : (uint32_t)blocks;
  return VFR_OK;
}

/* Blocks of the wave-queue kernel that fill the device: resident blocks per
 * CU for this kernel instance times the CUs. */
template <class K>
static uint32_t wq_blocks(K kernel) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 1;
  return (uint32_t)per_cu * (uint32_t)cus;
}

template <int PROJ, bool BIG, int TFL, bool PSIT>
static void launch_wq1(hipStream_t st, const TravArgs& a) {
  static const uint32_t blocks = wq_blocks(k_trav_wq<PROJ, BIG, TFL, PSIT>);
  hipLaunchKernelGGL((k_trav_wq<PROJ, BIG, TFL, PSIT>), dim3(blocks), dim3(256), 0, st, a);
}

template <bool PSIT>
static void launch_wq(uint32_t proj, bool big, bool tfl, hipStream_t st, const TravArgs& a) {
  if (big) launch_wq1<VFR_TRAV_MIXED, true, false, PSIT>(st, a);
  else if (proj == VFR_TRAV_ORTHO && tfl) launch_wq1<VFR_TRAV_ORTHO, false, true, PSIT>(st, a);
  else if (proj == VFR_TRAV_ORTHO) launch_wq1<VFR_TRAV_ORTHO, false, false, PSIT>(st, a);
  else if (proj == VFR_TRAV_PERSP && tfl) launch_wq1<VFR_TRAV_PERSP, false, true, PSIT>(st, a);
  else if (proj == VFR_TRAV_PERSP) launch_wq1<VFR_TRAV_PERSP, false, false, PSIT>(st, a);
  else if (tfl) launch_wq1<VFR_TRAV_MIXED, false, true, PSIT>(st, a);
  else launch_wq1<VFR_TRAV_MIXED, false, false, PSIT>(st, a);
}

static void launch_ftb_ldsr(uint32_t proj, bool tfl, dim3 grid, dim3 block, hipStream_t st, const TravArgs& a) {
  if (proj == VFR_TRAV_ORTHO && tfl)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_ORTHO, false, true, false, true>), grid, block, 0, st, a);
  else if (proj == VFR_TRAV_ORTHO)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_ORTHO, false, false, false, true>), grid, block, 0, st, a);
  else if (proj == VFR_TRAV_PERSP && tfl)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_PERSP, false, true, false, true>), grid, block, 0, st, a);
  else if (proj == VFR_TRAV_PERSP)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_PERSP, false, false, false, true>), grid, block, 0, st, a);
  else if (tfl)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_MIXED, false, true, false, true>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_MIXED, false, false, false, true>), grid, block, 0, st, a);
}

/* tfl: 0 the global TF table, 1 the launch's slots in LDS, 2 per-block slots */
template <bool PSIT, int TFL>
static void launch_ftb_t(uint32_t proj, dim3 grid, dim3 block, hipStream_t st, const TravArgs& a) {
  if (proj == VFR_TRAV_ORTHO)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_ORTHO, false, TFL, PSIT>), grid, block, 0, st, a);
  else if (proj == VFR_TRAV_PERSP)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_PERSP, false, TFL, PSIT>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_MIXED, false, TFL, PSIT>), grid, block, 0, st, a);
}

template <bool PSIT>
static void launch_ftb(uint32_t proj, int tfl, dim3 grid, dim3 block, hipStream_t st, const TravArgs& a) {
  if (tfl == 2) launch_ftb_t<PSIT, 2>(proj, grid, block, st, a);
  else if (tfl == 1) launch_ftb_t<PSIT, 1>(proj, grid, block, st, a);
  else launch_ftb_t<PSIT, 0>(proj, grid, block, st, a);
}

/* The largest float T with 1 - T >= term in float arithmetic (NaN for a NaN
 * term, which no T meets): the front-to-back walk's termination test as one
 * compare of its transparency. */
static float term_threshold(float term) {
  if (term != term) return term;
  float t = 1.0f - term;
  while (t < 2.0f && 1.0f - std::nextafter(t, 4.0f) >= term) t = std::nextafter(t, 4.0f);
  while (t > -2.0f && !(1.0f - t >= term)) t = std::nextafter(t, -4.0f);
  return t;
}

int vfr_k_traverse(uint32_t mode, uint32_t proj, uint64_t n_records, uint32_t occupancy, const TravArgs& a0,
                   uint32_t max_rays, int kernel, uint32_t grid_waves, void* stream) {
  if (max_rays == 0) return VFR_OK;
  /* per-block TF staging is the one-ray-per-thread kernel's */
  TravArgs a = a0;
  a.term_tr = term_threshold(a.term);
  if (kernel != VFR_WALK_THREAD) a.tf_block = 0;
  const bool lin = a.composite != VFR_COMPOSITE_PPI;
  if (kernel == VFR_WALK_QUEUE && mode == VFR_MODE_FTB && !lin) {
    const bool tfl = a.tf_lds_n >= 1 && a.tf_lds_n <= VFR_TF_LDS_SLOTS;
    const bool big = n_records >= VFR_BIG_RECORDS;
    if (a.psi_table)
      launch_wq<true>(proj, big, tfl, (hipStream_t)stream, a);
    else
      launch_wq<false>(proj, big, tfl, (hipStream_t)stream, a);
    CHECK_HIP(hipGetLastError());
    return VFR_OK;
  }
  if (kernel == VFR_WALK_REFILL && mode == VFR_MODE_FTB && !lin) {
    uint32_t blocks = (grid_waves + 3u) / 4u;
    if (blocks == 0) blocks = 1;
    const uint32_t burst = 8;
    if (occupancy >= 6)
      launch_persist<6>(blocks, stream, a, burst);
    else if (occupancy == 5)
      launch_persist<5>(blocks, stream, a, burst);
    else if (occupancy == 4)
      launch_persist<4>(blocks, stream, a, burst);
    else
      launch_persist<1>(blocks, stream, a, burst);
    CHECK_HIP(hipGetLastError());
    return VFR_OK;
  }
  dim3 grid((max_rays + 255) / 256), block(256);
  hipStream_t st = (hipStream_t)stream;
  /* TFs in LDS: the launch's slots when they fit, or per block */
  const bool tfl = a.tf_lds_n >= 1 && a.tf_lds_n <= VFR_TF_LDS_SLOTS;
  const int tfmode = tfl ? 1 : (a.tf_block ? 2 : 0);
  if (mode == VFR_MODE_REF && lin)
    hipLaunchKernelGGL((k_trav<VFR_MODE_REF_LINEAR, VFR_TRAV_MIXED, true, false, false>), grid, block, 0, st, a);
  else if (lin)
    /* linear compositing on the front-to-back walk: one general variant
     * (any projection, 64-bit record addressing, TF from global memory) */
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB_LINEAR, VFR_TRAV_MIXED, true, false, false>), grid, block, 0, st, a);
  else if (mode == VFR_MODE_REF)
    hipLaunchKernelGGL((k_trav<VFR_MODE_REF, VFR_TRAV_MIXED, true, false, false>), grid, block, 0, st, a);
  else if (n_records >= VFR_BIG_RECORDS && a.psi_table)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_MIXED, true, false, true>), grid, block, 0, st, a);
  else if (n_records >= VFR_BIG_RECORDS)
    hipLaunchKernelGGL((k_trav<VFR_MODE_FTB, VFR_TRAV_MIXED, true, false, false>), grid, block, 0, st, a);
  else if (a.psi_table)
    launch_ftb<true>(proj, tfmode, grid, block, st, a);
  else if (a.lds_records)
    launch_ftb_ldsr(proj, tfl, grid, block, st, a);
  else
    launch_ftb<false>(proj, tfmode, grid, block, st, a);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* ---------------------------------------------------------- preprocessing */

/* Face keys: for slot idx = t*4+f the face's vertex ids (tet[(f+k)%4],
 * k = 0..2, vfRayPreComp.cc:442-446) sorted a < b < c; key = (a << bits) | b,
 * c kept aside.  Flags tets with a repeated vertex id. */
__global__ void k_face_keys(const uint32_t* __restrict__ tets, uint32_t nT, uint32_t bits,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                            uint32_t* __restrict__ cval, uint32_t* __restrict__ bad) {
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)nT * 4) return;
  uint32_t t = (uint32_t)(idx >> 2), f = (uint32_t)(idx & 3);
  uint32_t a = tets[t * 4 + ((f + 0) & 3)], b = tets[t * 4 + ((f + 1) & 3)],
           c = tets[t * 4 + ((f + 2) & 3)], d = tets[t * 4 + ((f + 3) & 3)];
  if (f == 0 && (a == b || a == c || a == d || b == c || b == d || c == d)) atomicAdd(bad, 1u);
  uint32_t x = min(a, min(b, c)), z = max(a, max(b, c)), y = a ^ b ^ c ^ x ^ z;
  keys[idx] = ((uint64_t)x << bits) | y;
  vals[idx] = (uint32_t)idx;
  cval[idx] = z;
}

/* Adjacent match over the sorted keys: within the run of equal (a,b), the
 * slots with the same c share the face; the neighbour is the smallest other
 * tet id among them (buildCon's ascending search, vfRayPreComp.cc:450-490),
 * else the tet itself (:448). */
__global__ void k_match(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                        const uint32_t* __restrict__ cval, uint64_t n, uint32_t* __restrict__ con) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  const uint32_t idx = vals[i], c = cval[idx], t = idx >> 2;
  uint32_t best = 0xFFFFFFFFu;
  for (uint64_t j = i; j > 0 && keys[j - 1] == key; --j) {
    uint32_t o = vals[j - 1];
    if (cval[o] == c && (o >> 2) != t) best = min(best, o >> 2);
  }
  for (uint64_t j = i + 1; j < n && keys[j] == key; ++j) {
    uint32_t o = vals[j];
    if (cval[o] == c && (o >> 2) != t) best = min(best, o >> 2);
  }
  con[idx] = best == 0xFFFFFFFFu ? t : best;
}

__global__ void k_ext_flags(const uint32_t* __restrict__ con, uint64_t n, uint32_t* __restrict__ flag) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = con[i] == (uint32_t)(i >> 2) ? 1u : 0u;
}

__global__ void k_ext_write(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                            uint64_t n, uint32_t* __restrict__ extf) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  uint32_t p = pos[i];
  extf[p * 2] = (uint32_t)(i >> 2);
  extf[p * 2 + 1] = (uint32_t)(i & 3);
}

DEV uint32_t spread10(uint32_t x) {
  x &= 0x3FFu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

DEV uint32_t morton3(float x, float y, float z, const float* bb) {
  float sx = (x - bb[0]) * bb[3], sy = (y - bb[1]) * bb[4], sz = (z - bb[2]) * bb[5];
  uint32_t ix = (uint32_t)fminf(fmaxf(sx, 0.0f), 1023.0f);
  uint32_t iy = (uint32_t)fminf(fmaxf(sy, 0.0f), 1023.0f);
  uint32_t iz = (uint32_t)fminf(fmaxf(sz, 0.0f), 1023.0f);
  return spread10(ix) | (spread10(iy) << 1) | (spread10(iz) << 2);
}

/* Vertex order key: the first position, in the new tet order, of a tet that
 * holds the vertex (0xFFFFFFFF for a vertex no tet holds; key pre-filled). */
__global__ void k_vert_first(const uint32_t* __restrict__ tets, uint32_t nT, const uint32_t* __restrict__ tinv,
                             uint32_t* __restrict__ key) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nT) return;
  const uint32_t t = tinv[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) atomicMin(key + tets[4 * i + k], t);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_tet_morton(const float4* __restrict__ v, const uint32_t* __restrict__ tets,
                             uint32_t nT, const float* bb, uint32_t* __restrict__ code,
                             uint32_t* __restrict__ id) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nT) return;
  float4 a = v[tets[i * 4]], b = v[tets[i * 4 + 1]], c = v[tets[i * 4 + 2]], d = v[tets[i * 4 + 3]];
  code[i] = morton3((a.x + b.x + c.x + d.x) * 0.25f, (a.y + b.y + c.y + d.y) * 0.25f,
                    (a.z + b.z + c.z + d.z) * 0.25f, bb);
  id[i] = i;
}

__global__ void k_invert(const uint32_t* __restrict__ perm, uint32_t n, uint32_t* __restrict__ inv) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[perm[i]] = i;
}

__global__ void k_gather_vert(const float4* __restrict__ v, const uint32_t* __restrict__ vperm,
                              uint32_t nV, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nV) out[i] = v[vperm[i]];
}

/* Internal tet records: vertex ids through the vertex permutation (tv);
 * per face the neighbour word (internal id << 2 | the face of the neighbour
 * that links back, the back-lookup of vfRayKernelCPU.cc:359-366) and the
 * vertex word: the neighbour's vertex off that face (the tet's own vertex
 * opposite the face when there is no linked neighbour), so a walk can fetch
 * the next tet's fourth vertex without waiting for its record, and the
 * neighbour's slots of the face's vertices. */
__global__ void k_build_rec(const uint32_t* __restrict__ tets, const uint32_t* __restrict__ con,
                            const uint32_t* __restrict__ perm, const uint32_t* __restrict__ tinv,
                            const uint32_t* __restrict__ vinv, uint32_t nT,
                            TetRec* __restrict__ rec, uint4* __restrict__ tv) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nT) return;
  uint32_t t = perm[i];
  TetRec r;
  uint32_t v[4], n[4], o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = vinv[tets[t * 4 + k]];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    uint32_t u = con[t * 4 + f];
    o[f] = v[(f + 3) & 3];
    if (u == t) { n[f] = VFR_NB_EXT; continue; }
    uint32_t kk = 4;
    for (uint32_t q = 0; q < 4; ++q)
      if (con[u * 4 + q] == t) { kk = q; break; }
    n[f] = kk == 4 ? VFR_NB_BROKEN : ((tinv[u] << 2) | kk);
    if (kk != 4) {
      o[f] = vinv[tets[u * 4 + ((kk + 3) & 3)]];
      /* the slot map across the face, byte i = the image of slot i: each
       * vertex of the face to its slot in the neighbour, the vertex off the
       * face to the neighbour's vertex off the face */
      uint32_t map = ((kk + 3u) & 3u) << (8u * ((f + 3u) & 3u));
      for (uint32_t p = 0; p < 3; ++p) {
        const uint32_t id = tets[t * 4 + ((f + p) & 3)];
        uint32_t q = 0;
        while (q < 3 && tets[u * 4 + q] != id) ++q;
        map |= q << (8u * ((f + p) & 3u));
      }
      uint32_t pi = 0;
      while (pi < VFR_N_SLOT_PERMS && VFR_SLOT_PERMS[pi] != map) ++pi;
      /* not a permutation: the neighbour's linked face is not this face */
      if (pi == VFR_N_SLOT_PERMS) n[f] = VFR_NB_BROKEN;
      else o[f] |= pi << VFR_REC_VERT_BITS;
    }
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    r.f[2 * f] = n[f];
    r.f[2 * f + 1] = o[f];
  }
  rec[i] = r;
  tv[i] = make_uint4(v[0], v[1], v[2], v[3]);
}

/* Adds `add` to the tet index of every linked neighbour word of records
 * [0, n): a mesh's records copied into an arena at record `add`. */
__global__ void k_rebase_links(TetRec* __restrict__ rec, uint32_t n, uint32_t add) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t w = rec[i].f[2 * k];
    if (w < VFR_NB_BROKEN) rec[i].f[2 * k] = w + (add << 2);
  }
}

int vfr_k_rebase_links(TetRec* rec, uint32_t n, uint32_t add, void* stream) {
  if (n == 0 || add == 0) return VFR_OK;
  hipLaunchKernelGGL(k_rebase_links, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, rec, n, add);
  CHECK_HIP(hipGetLastError());
  return VFR_OK;
}

/* ------------------------------------------------- rotate + renormalize */

/* Order-preserving integer key of a float (-0 counts as +0), so that the
 * min/max of the keys are the keys of the min/max under `<`. */
DEV uint32_t fkey_ord(float f) {
  uint32_t u = __float_as_uint(f);
  if (u == 0x80000000u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static inline float fkey_dec(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct Mat16 { float m[16]; };

__global__ void k_minmax_init(uint32_t* mm) {
  if (threadIdx.x < 8) mm[threadIdx.x] = threadIdx.x < 4 ? 0xFFFFFFFFu : 0u;
}

/* rotateVolume (vfRayPreComp.cc:886-906): row r of the result is
 * sum_c mv[r*4+c] * (x, y, z, 1)[c], accumulated from 0 in c order; rows 0-2
 * replace x, y, z and the scalar is kept.  With `rot` clear the vertices are
 * only scanned.  mm receives the keys of the componentwise min (0..3) and max
 * (4..7) of the result. */
__global__ __launch_bounds__(256) void k_rotate_verts(float4* __restrict__ v, uint32_t n, Mat16 M, int rot,
                                                      uint32_t* __restrict__ mm) {
  uint32_t lo[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[4] = {0u, 0u, 0u, 0u};
  /* grid-stride: a few thousand blocks each reduce many vertices, so the
   * eight global atomics run once per block */
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 p = v[i];
    if (rot) {
      const float o[4] = {p.x, p.y, p.z, 1.0f};
      float r[3];
      for (int a = 0; a < 3; ++a) {
        float acc = 0.0f;
        for (int c = 0; c < 4; ++c) acc = acc + M.m[a * 4 + c] * o[c];
        r[a] = acc;
      }
      p.x = r[0]; p.y = r[1]; p.z = r[2];
      v[i] = p;
    }
    const float q[4] = {p.x, p.y, p.z, p.w};
    for (int a = 0; a < 4; ++a) {
      lo[a] = min(lo[a], fkey_ord(q[a]));
      hi[a] = max(hi[a], fkey_ord(q[a]));
    }
  }
  for (int a = 0; a < 4; ++a)
    for (int o = 32; o > 0; o >>= 1) {
      lo[a] = min(lo[a], (uint32_t)__shfl_xor((int)lo[a], o, 64));
      hi[a] = max(hi[a], (uint32_t)__shfl_xor((int)hi[a], o, 64));
    }
  __shared__ uint32_t blo[4][4], bhi[4][4];
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0)
    for (int a = 0; a < 4; ++a) { blo[w][a] = lo[a]; bhi[w][a] = hi[a]; }
  __syncthreads();
  if (threadIdx.x < 4) {
    const uint32_t a = threadIdx.x;
    uint32_t l = blo[0][a], h = bhi[0][a];
    for (uint32_t k = 1; k < (blockDim.x >> 6); ++k) { l = min(l, blo[k][a]); h = max(h, bhi[k][a]); }
    atomicMin(&mm[a], l);
    atomicMax(&mm[4 + a], h);
  }
}

/* normalizeVertices' update (vfRayPreComp.cc:136-143). */
__global__ __launch_bounds__(256) void k_normalize_verts(float4* __restrict__ v, uint32_t n, float c0, float c1,
                                                         float c2, float sc, float s0, float ss) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = v[i];
  p.x = (p.x - c0) * sc;
  p.y = (p.y - c1) * sc;
  p.z = (p.z - c2) * sc;
  p.w = (p.w - s0) * ss;
  v[i] = p;
}

int vfr_k_rotate_normalize(float4* v, uint32_t n, const float* mv, int normalize, float lo[3], float hi[3],
                           void* stream) {
  if (n == 0) return VFR_OK;
  hipStream_t st = (hipStream_t)stream;
  uint32_t* mm = nullptr;
  CHECK_HIP(hipMalloc(&mm, 8 * sizeof(uint32_t)));
  Mat16 M;
  for (int k = 0; k < 16; ++k) M.m[k] = mv ? mv[k] : 0.0f;
  uint32_t h[8];
  hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, st, mm);
  const uint32_t rb = (n + 255) / 256 < 2048u ? (n + 255) / 256 : 2048u;
  hipLaunchKernelGGL(k_rotate_verts, dim3(rb), dim3(256), 0, st, v, n, M, mv ? 1 : 0, mm);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(mm);
  CHECK_HIP(e);
  float mn[4], mx[4];
  for (int a = 0; a < 4; ++a) { mn[a] = fkey_dec(h[a]); mx[a] = fkey_dec(h[4 + a]); }
  if (normalize) {
    float center[3], sc, ss;
    vfr_normalize_params(mn, mx, center, &sc, &ss);
    hipLaunchKernelGGL(k_normalize_verts, dim3((n + 255) / 256), dim3(256), 0, st, v, n, center[0], center[1],
                       center[2], sc, mn[3], ss);
    CHECK_HIP(hipGetLastError());
    CHECK_HIP(hipStreamSynchronize(st));
    /* the update is monotonic in each coordinate: the extremes map to the
       extremes */
    for (int a = 0; a < 3; ++a) {
      lo[a] = (mn[a] - center[a]) * sc;
      hi[a] = (mx[a] - center[a]) * sc;
    }
  } else {
    for (int a = 0; a < 3; ++a) { lo[a] = mn[a]; hi[a] = mx[a]; }
  }
  return VFR_OK;
}

/* Ids out of range: `first` receives the smallest index i with
 * a[i] >= limit, or, with `faces` set, a[i] > 3 at odd i (the face numbers of
 * (tet, face) pairs). */
__global__ void k_check_ids(const uint32_t* __restrict__ a, uint64_t n, uint32_t limit, int faces,
                            unsigned long long* __restrict__ first) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t lim = (faces && (i & 1u)) ? 4u : limit;
  if (a[i] >= lim) atomicMin(first, (unsigned long long)i);
}

/* buildLmt's longest edge (vfRayPreComp.cc:808-833): per tet the six edges
 * (0,1) (0,2) (0,3) (1,2) (1,3) (2,3), length = sqrt in float of
 * (dx*dx + dy*dy) + dz*dz, unfused; the maximum over every tet as the bits of
 * a non-negative float (NaN lengths are skipped, as the reference's `>` test
 * skips them). */
__global__ __launch_bounds__(256) void k_max_edge(const float4* __restrict__ v, const uint4* __restrict__ tets,
                                                  uint32_t nT, uint32_t* __restrict__ out) {
  float m = 0.0f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nT; i += gridDim.x * blockDim.x) {
    const uint4 q = tets[i];
    const float4 p[4] = {v[q.x], v[q.y], v[q.z], v[q.w]};
    const int E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const float4 a = p[E[j][0]], b = p[E[j][1]];
      const float e0 = a.x - b.x, e1 = a.y - b.y, e2 = a.z - b.z;
      const float len = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
      if (len > m) m = len;
    }
  }
  uint32_t bits = __float_as_uint(m);
  for (int o = 32; o > 0; o >>= 1) bits = max(bits, (uint32_t)__shfl_xor((int)bits, o, 64));
  __shared__ uint32_t wb[4];
  if ((threadIdx.x & 63u) == 0) wb[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t k = 1; k < (blockDim.x >> 6); ++k) bits = max(bits, wb[k]);
    if (bits) atomicMax(out, bits);
  }
}

/* Smallest index of a vertex whose z is +0 or -0. */
__global__ void k_first_zero_z(const float4* __restrict__ v, uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && v[i].z == 0.0f) atomicMin(out, i);
}

__global__ void k_ext_internal(const uint32_t* __restrict__ extf, uint32_t n,
                               const uint32_t* __restrict__ tinv, uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i * 2] = tinv[extf[i * 2]];
  out[i * 2 + 1] = extf[i * 2 + 1];
}

/* Vertex ids of each external face's slots f, f+1, f+2 and the one off it,
 * from its (internal tet, face) and the tet's ids. */
__global__ void k_face_vids(const uint32_t* __restrict__ ext_int, const uint4* __restrict__ tv, uint32_t n,
                            uint4* __restrict__ fvid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = ext_int[2 * i], f = ext_int[2 * i + 1];
  const uint4 rv = tv[t];
  fvid[i] = make_uint4(sel4(rv, f), sel4(rv, (f + 1u) & 3u), sel4(rv, (f + 2u) & 3u), sel4(rv, (f + 3u) & 3u));
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 16); }
  template <class T> T* as() const { return static_cast<T*>(p); }
  void* release() { void* q = p; p = nullptr; return q; }
};

inline unsigned nblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

#define ALLOC(b, n) CHECK_HIP((b).alloc(n))

/* Builds every device structure of a mesh.  h_xyzs (already normalized if
 * requested) and h_tets are caller order; h_con/h_extf optional (the .con /
 * .extf caches). */
int vfr_dev_build_mesh(DevMesh& m, const float* h_xyzs, const uint32_t* h_tets,
                       const uint32_t* h_con, const uint32_t* h_extf, uint32_t n_ext_given,
                       bool normalize, bool morton, const float* h_lmt, float bb_lo[3], float bb_hi[3],
                       void* stream_v) {
  hipStream_t st = (hipStream_t)stream_v;
  const uint32_t nV = m.n_verts, nT = m.n_tets;
  const uint64_t nF = (uint64_t)nT * 4;
  DevBuf d_xyzs, d_tets, d_con, d_bad, d_first;
  ALLOC(d_xyzs, sizeof(float) * 4 * (size_t)nV);
  ALLOC(d_tets, sizeof(uint32_t) * 4 * (size_t)nT);
  ALLOC(d_con, sizeof(uint32_t) * 4 * (size_t)nT);
  ALLOC(d_bad, sizeof(uint32_t) * 4);
  ALLOC(d_first, sizeof(unsigned long long) * 3);
  CHECK_HIP(hipMemcpyAsync(d_xyzs.p, h_xyzs, sizeof(float) * 4 * (size_t)nV, hipMemcpyHostToDevice, st));
  CHECK_HIP(hipMemcpyAsync(d_tets.p, h_tets, sizeof(uint32_t) * 4 * (size_t)nT, hipMemcpyHostToDevice, st));
  CHECK_HIP(hipMemsetAsync(d_bad.p, 0, 16, st));
  CHECK_HIP(hipMemsetAsync(d_first.p, 0xFF, sizeof(unsigned long long) * 3, st));
  /* ids in range: tets' vertices, given neighbours, given external faces */
  DevBuf d_extf;
  if (h_extf) ALLOC(d_extf, sizeof(uint32_t) * 2 * (size_t)(n_ext_given ? n_ext_given : 1));
  unsigned long long* first = d_first.as<unsigned long long>();
  hipLaunchKernelGGL(k_check_ids, dim3(nblk(nF)), dim3(256), 0, st, d_tets.as<uint32_t>(), nF, nV, 0, first);
  if (h_con) {
    CHECK_HIP(hipMemcpyAsync(d_con.p, h_con, sizeof(uint32_t) * nF, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_check_ids, dim3(nblk(nF)), dim3(256), 0, st, d_con.as<uint32_t>(), nF, nT, 0, first + 1);
  }
  if (h_extf && n_ext_given) {
    CHECK_HIP(hipMemcpyAsync(d_extf.p, h_extf, sizeof(uint32_t) * 2 * (size_t)n_ext_given,
                             hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_check_ids, dim3(nblk(2ull * n_ext_given)), dim3(256), 0, st, d_extf.as<uint32_t>(),
                       2ull * n_ext_given, nT, 1, first + 2);
  }
  CHECK_HIP(hipGetLastError());
  unsigned long long hf[3];
  CHECK_HIP(hipMemcpyAsync(hf, first, sizeof(hf), hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipStreamSynchronize(st));
  if (hf[0] != ~0ull) {
    vfr_set_error("tet %llu references vertex %u >= %u", hf[0] / 4, h_tets[hf[0]], nV);
    return VFR_EMESH;
  }
  if (hf[1] != ~0ull) {
    vfr_set_error("con entry %llu = %u >= %u", hf[1], h_con[hf[1]], nT);
    return VFR_EMESH;
  }
  if (hf[2] != ~0ull) {
    vfr_set_error("extf entry %llu invalid", hf[2] / 2);
    return VFR_EMESH;
  }
  /* normalizeVertices on the device; the bounds of the result */
  {
    const int rc = vfr_k_rotate_normalize(d_xyzs.as<float4>(), nV, nullptr, normalize ? 1 : 0, bb_lo, bb_hi, st);
    if (rc != VFR_OK) return rc;
  }
  /* buildLmt on the device unless the limits are given: the longest edge,
   * and the z range starting from maxZ = -1, minZ = 1 (vfRayPreComp.cc:
   * 835-850).  An extreme z of zero is the first zero in vertex order, as the
   * reference's strict comparisons keep it, with its sign. */
  if (h_lmt) {
    memcpy(m.lmt, h_lmt, sizeof(float) * 3);
  } else {
    DevBuf d_lm;
    ALLOC(d_lm, sizeof(uint32_t) * 2);
    uint32_t hl[2] = {0u, 0xFFFFFFFFu};
    CHECK_HIP(hipMemcpyAsync(d_lm.p, hl, sizeof(hl), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_max_edge, dim3(nblk(nT) < 4096u ? nblk(nT) : 4096u), dim3(256), 0, st, d_xyzs.as<float4>(),
                       d_tets.as<uint4>(), nT, d_lm.as<uint32_t>());
    const bool zero = bb_hi[2] == 0.0f || bb_lo[2] == 0.0f;
    if (zero)
      hipLaunchKernelGGL(k_first_zero_z, dim3(nblk(nV)), dim3(256), 0, st, d_xyzs.as<float4>(), nV,
                         d_lm.as<uint32_t>() + 1);
    CHECK_HIP(hipGetLastError());
    CHECK_HIP(hipMemcpyAsync(hl, d_lm.p, sizeof(hl), hipMemcpyDeviceToHost, st));
    CHECK_HIP(hipStreamSynchronize(st));
    float z0 = 0.0f;
    if (zero && hl[1] < nV)
      CHECK_HIP(hipMemcpy(&z0, d_xyzs.as<float>() + 4 * (size_t)hl[1] + 2, sizeof(float), hipMemcpyDeviceToHost));
    memcpy(&m.lmt[0], &hl[0], sizeof(float));
    m.lmt[1] = bb_hi[2] > -1.0f ? (bb_hi[2] == 0.0f ? z0 : bb_hi[2]) : -1.0f;
    m.lmt[2] = bb_lo[2] < 1.0f ? (bb_lo[2] == 0.0f ? z0 : bb_lo[2]) : 1.0f;
  }
  uint32_t n_ext = 0;
  if (!h_con) {
    /* sort-and-match */
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)nV) bits++;
    DevBuf k0, k1, v0, v1, cv, tmp;
    ALLOC(k0, 8 * nF); ALLOC(k1, 8 * nF); ALLOC(v0, 4 * nF); ALLOC(v1, 4 * nF); ALLOC(cv, 4 * nF);
    hipLaunchKernelGGL(k_face_keys, dim3(nblk(nF)), dim3(256), 0, st, d_tets.as<uint32_t>(), nT,
                       bits, k0.as<uint64_t>(), v0.as<uint32_t>(), cv.as<uint32_t>(),
                       d_bad.as<uint32_t>());
    CHECK_HIP(hipGetLastError());
    size_t tb = 0;
    CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0.as<uint64_t>(), k1.as<uint64_t>(),
                                                 v0.as<uint32_t>(), v1.as<uint32_t>(), (int)nF, 0,
                                                 (int)(2 * bits), st));
    ALLOC(tmp, tb);
    CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.as<uint64_t>(), k1.as<uint64_t>(),
                                                 v0.as<uint32_t>(), v1.as<uint32_t>(), (int)nF, 0,
                                                 (int)(2 * bits), st));
    hipLaunchKernelGGL(k_match, dim3(nblk(nF)), dim3(256), 0, st, k1.as<uint64_t>(),
                       v1.as<uint32_t>(), cv.as<uint32_t>(), nF, d_con.as<uint32_t>());
    CHECK_HIP(hipGetLastError());
  }
  uint32_t bad = 0;
  CHECK_HIP(hipMemcpyAsync(&bad, d_bad.p, 4, hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipStreamSynchronize(st));
  if (bad) {
    vfr_set_error("mesh has %u tetrahedra with a repeated vertex id", bad);
    return VFR_EMESH;
  }
  if (h_extf) {
    n_ext = n_ext_given;
  } else {
    DevBuf flag, pos, tmp;
    ALLOC(flag, 4 * nF); ALLOC(pos, 4 * nF);
    hipLaunchKernelGGL(k_ext_flags, dim3(nblk(nF)), dim3(256), 0, st, d_con.as<uint32_t>(), nF,
                       flag.as<uint32_t>());
    CHECK_HIP(hipGetLastError());
    size_t tb = 0;
    CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.as<uint32_t>(), pos.as<uint32_t>(),
                                               (int)nF, st));
    ALLOC(tmp, tb);
    CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.as<uint32_t>(), pos.as<uint32_t>(),
                                               (int)nF, st));
    uint32_t lastp = 0, lastf = 0;
    CHECK_HIP(hipMemcpyAsync(&lastp, pos.as<uint32_t>() + nF - 1, 4, hipMemcpyDeviceToHost, st));
    CHECK_HIP(hipMemcpyAsync(&lastf, flag.as<uint32_t>() + nF - 1, 4, hipMemcpyDeviceToHost, st));
    CHECK_HIP(hipStreamSynchronize(st));
    n_ext = lastp + lastf;
    ALLOC(d_extf, sizeof(uint32_t) * 2 * (size_t)n_ext);
    hipLaunchKernelGGL(k_ext_write, dim3(nblk(nF)), dim3(256), 0, st, flag.as<uint32_t>(),
                       pos.as<uint32_t>(), nF, d_extf.as<uint32_t>());
    CHECK_HIP(hipGetLastError());
  }
  /* orderings */
  DevBuf perm, tinv, vperm, vinv;
  ALLOC(perm, 4 * (size_t)nT); ALLOC(tinv, 4 * (size_t)nT);
  ALLOC(vperm, 4 * (size_t)nV); ALLOC(vinv, 4 * (size_t)nV);
  if (morton) {
    float hb[6];
    for (int a = 0; a < 3; ++a) {
      hb[a] = bb_lo[a];
      hb[3 + a] = bb_hi[a] > bb_lo[a] ? 1023.0f / (bb_hi[a] - bb_lo[a]) : 0.0f;
    }
    DevBuf bb, code0, code1, id0, tmp;
    ALLOC(bb, sizeof(hb));
    CHECK_HIP(hipMemcpyAsync(bb.p, hb, sizeof(hb), hipMemcpyHostToDevice, st));
    uint32_t nmax = nT > nV ? nT : nV;
    ALLOC(code0, 4 * (size_t)nmax); ALLOC(code1, 4 * (size_t)nmax); ALLOC(id0, 4 * (size_t)nmax);
    size_t tb = 0, tb2 = 0;
    CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, code0.as<uint32_t>(), code1.as<uint32_t>(),
                                                 id0.as<uint32_t>(), perm.as<uint32_t>(), (int)nT, 0, 30, st));
    CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, code0.as<uint32_t>(), code1.as<uint32_t>(),
                                                 id0.as<uint32_t>(), vperm.as<uint32_t>(), (int)nV, 0, 32, st));
    ALLOC(tmp, tb > tb2 ? tb : tb2);
    /* tets by the Morton code of their centroid; vertices in the order the
     * sorted tets first reach them, so the vertices a wave's rays fetch
     * together sit together as their tets do (walk 1 % faster than a
     * Morton order of the vertices themselves) */
    hipLaunchKernelGGL(k_tet_morton, dim3(nblk(nT)), dim3(256), 0, st, d_xyzs.as<float4>(),
                       d_tets.as<uint32_t>(), nT, bb.as<float>(), code0.as<uint32_t>(),
                       id0.as<uint32_t>());
    CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, code0.as<uint32_t>(), code1.as<uint32_t>(),
                                                 id0.as<uint32_t>(), perm.as<uint32_t>(), (int)nT, 0, 30, st));
    hipLaunchKernelGGL(k_invert, dim3(nblk(nT)), dim3(256), 0, st, perm.as<uint32_t>(), nT, tinv.as<uint32_t>());
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(nV)), dim3(256), 0, st, code0.as<uint32_t>(), nV, 0xFFFFFFFFu);
    hipLaunchKernelGGL(k_iota, dim3(nblk(nV)), dim3(256), 0, st, id0.as<uint32_t>(), nV);
    hipLaunchKernelGGL(k_vert_first, dim3(nblk(nT)), dim3(256), 0, st, d_tets.as<uint32_t>(), nT,
                       tinv.as<uint32_t>(), code0.as<uint32_t>());
    CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb2, code0.as<uint32_t>(), code1.as<uint32_t>(),
                                                 id0.as<uint32_t>(), vperm.as<uint32_t>(), (int)nV, 0, 32, st));
    CHECK_HIP(hipStreamSynchronize(st));
  } else {
    hipLaunchKernelGGL(k_iota, dim3(nblk(nT)), dim3(256), 0, st, perm.as<uint32_t>(), nT);
    hipLaunchKernelGGL(k_iota, dim3(nblk(nV)), dim3(256), 0, st, vperm.as<uint32_t>(), nV);
  }
  hipLaunchKernelGGL(k_invert, dim3(nblk(nT)), dim3(256), 0, st, perm.as<uint32_t>(), nT, tinv.as<uint32_t>());
  hipLaunchKernelGGL(k_invert, dim3(nblk(nV)), dim3(256), 0, st, vperm.as<uint32_t>(), nV, vinv.as<uint32_t>());
  DevBuf vert, rec, tv, ext_int, fvid;
  ALLOC(vert, sizeof(float) * 4 * (size_t)nV);
  ALLOC(rec, sizeof(TetRec) * (size_t)nT);
  ALLOC(tv, sizeof(uint4) * (size_t)nT);
  ALLOC(ext_int, sizeof(uint32_t) * 2 * (size_t)n_ext);
  ALLOC(fvid, sizeof(uint4) * (size_t)n_ext);
  hipLaunchKernelGGL(k_gather_vert, dim3(nblk(nV)), dim3(256), 0, st, d_xyzs.as<float4>(),
                     vperm.as<uint32_t>(), nV, vert.as<float4>());
  hipLaunchKernelGGL(k_build_rec, dim3(nblk(nT)), dim3(256), 0, st, d_tets.as<uint32_t>(),
                     d_con.as<uint32_t>(), perm.as<uint32_t>(), tinv.as<uint32_t>(),
                     vinv.as<uint32_t>(), nT, rec.as<TetRec>(), tv.as<uint4>());
  if (n_ext) {
    hipLaunchKernelGGL(k_ext_internal, dim3(nblk(n_ext)), dim3(256), 0, st, d_extf.as<uint32_t>(),
                       n_ext, tinv.as<uint32_t>(), ext_int.as<uint32_t>());
    hipLaunchKernelGGL(k_face_vids, dim3(nblk(n_ext)), dim3(256), 0, st, ext_int.as<uint32_t>(),
                       tv.as<uint4>(), n_ext, fvid.as<uint4>());
  }
  CHECK_HIP(hipGetLastError());
  CHECK_HIP(hipStreamSynchronize(st));
  m.n_ext = n_ext;
  m.vert = vert.as<float>(); vert.release();
  m.rec = rec.as<TetRec>(); rec.release();
  m.tv = tv.as<uint32_t>(); tv.release();
  m.ext_int = ext_int.as<uint32_t>(); ext_int.release();
  m.fvid = fvid.as<uint32_t>(); fvid.release();
  m.extf = d_extf.as<uint32_t>(); d_extf.release();
  m.perm = perm.as<uint32_t>(); perm.release();
  m.vperm = vperm.as<uint32_t>(); vperm.release();
  m.con = d_con.as<uint32_t>(); d_con.release();
  m.bytes = sizeof(float) * 4ull * nV + (sizeof(TetRec) + 16ull) * (uint64_t)nT + 32ull * n_ext +
            4ull * nT + 4ull * nV + 16ull * nT;
  return VFR_OK;
}
