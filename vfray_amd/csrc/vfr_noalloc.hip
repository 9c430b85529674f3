/*
 * vfr_noalloc.hip -- vfr_render_noalloc: the render of vfr_render without
 * any device (or host heap) allocation.  Its per-pixel and per-view buffers
 * live in one file-scope device array, `scratch`, of VFR_SCRATCH_BYTES; its
 * device address is taken once per device when a context is created (the
 * library's init call) and kept per device.  Every call on one device shares
 * that device's array: calls that overlap on one device race, and each call
 * clears what must start at zero.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

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

#define CHECK_RC(x)                \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != VFR_OK) return rc_; \
  } while (0)

/* The scratch area of vfr_render_noalloc (one copy per device). */
static __device__ __align__(256) unsigned char scratch[VFR_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

/* Takes the device address of `scratch` on `device`, which must be the
 * current device; once per device. */
int vfr_noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    vfr_set_error("device ordinal %d beyond the %d the scratch table holds", device, kMaxDevices);
    return VFR_ENODEV;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return VFR_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    vfr_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return VFR_EHIP;
  }
  g_scratch_ptr[device] = p;
  return VFR_OK;
}

static void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

namespace {

inline uint64_t up256(uint64_t x) { return (x + 255u) & ~255ull; }

/* Byte offsets of the buffers in the scratch area; the entry keys come last,
 * so that everything before them does not depend on the slot count. */
struct Layout {
  uint64_t views, counters, stats, cnt, rays, rgba, steps, entry, pverts, entries, end;
};

Layout layout(uint32_t n_views, uint64_t pix, uint64_t n_verts, uint32_t kmax) {
  Layout L;
  uint64_t o = 0;
  L.views = o;    o = up256(o + sizeof(DevView) * n_views);
  L.counters = o; o = up256(o + 4ull * (n_views + 4));
  L.stats = o;    o = up256(o + 8ull * ST_N);
  L.cnt = o;      o = up256(o + 4ull * pix);
  L.rays = o;     o = up256(o + 8ull * pix);
  L.rgba = o;     o = up256(o + 16ull * pix);
  L.steps = o;    o = up256(o + 4ull * pix);
  L.entry = o;    o = up256(o + 4ull * pix);
  L.pverts = o;   o = up256(o + 16ull * n_views * n_verts);
  L.entries = o;  o = up256(o + 8ull * kmax * pix);
  L.end = o;
  return L;
}

template <class T>
T* at(void* base, uint64_t off) {
  return reinterpret_cast<T*>(static_cast<unsigned char*>(base) + off);
}

}  // namespace

extern "C" int vfr_render_noalloc(vfr_ctx* c, const vfr_mesh* m, const vfr_view* views, int n_views,
                                  const vfr_render_opts* opts, vfr_frame_out* outs, vfr_stats* stats) {
  if (!c || !m || !views || n_views < 1) { vfr_set_error("vfr_render_noalloc: bad argument"); return VFR_EINVAL; }
  if (vfr_mesh_ctx(m) != c) { vfr_set_error("mesh belongs to another context"); return VFR_EINVAL; }
  vfr_render_opts o{};
  if (opts) o = *opts;
  if (o.mode > VFR_MODE_FTB || o.composite > VFR_COMPOSITE_LINEAR_OPACITY) {
    vfr_set_error("vfr_render_noalloc: bad mode or compositing");
    return VFR_EINVAL;
  }
  if (o.use_graph || (o.flags & VFR_RENDER_PIPELINED)) {
    vfr_set_error("vfr_render_noalloc renders one batched launch set: no graph, no pipelined chunks");
    return VFR_EINVAL;
  }
  if ((o.flags & VFR_RENDER_PERSISTENT) && (o.flags & VFR_RENDER_WAVE_QUEUE)) {
    vfr_set_error("VFR_RENDER_PERSISTENT and VFR_RENDER_WAVE_QUEUE exclude each other");
    return VFR_EINVAL;
  }
  if (o.max_entries != 0 && o.max_entries != 8 && o.max_entries != 16 && o.max_entries != VFR_K_MAX) {
    vfr_set_error("max_entries must be 0 (chosen per call), 8, 16 or %u", VFR_K_MAX);
    return VFR_EINVAL;
  }
  if (!(o.term_alpha > 0.0f)) o.term_alpha = 0.99f;
  if (o.term_alpha > 1.0f) { vfr_set_error("term_alpha %g > 1", (double)o.term_alpha); return VFR_EINVAL; }
  if (o.waves_per_cu == 0) o.waves_per_cu = 16;
  if (o.refill == 0) o.refill = 32;
  if (o.waves_per_cu > 64 || o.refill > 64) { vfr_set_error("waves_per_cu / refill out of range"); return VFR_EINVAL; }
  for (int i = 0; i < n_views; ++i)
    if (outs && outs[i].rgb8 && !outs[i].rgba) {
      vfr_set_error("vfr_render_noalloc: view %d asks for rgb8 without rgba (the bytes come from it)", i);
      return VFR_EINVAL;
    }
  const int dev = vfr_ctx_device(c);
  CHECK_HIP(hipSetDevice(dev));
  void* base = scratch_ptr(dev);
  if (!base) { vfr_set_error("scratch area of device %d not set up", dev); return VFR_ESTATE; }
  const DevMesh* d = vfr_mesh_dev(m);
  float lo[3], hi[3];
  vfr_mesh_bbox(m, lo, hi);
  hipStream_t st = (hipStream_t)vfr_ctx_stream(c);
  const auto t0 = std::chrono::steady_clock::now();

  /* views: set up one at a time on the stack, their offsets as a plan's */
  uint64_t pix = 0, tiles = 0;
  uint32_t persp = 0, tf_lo = UINT32_MAX, tf_hi = 0;
  for (int i = 0; i < n_views; ++i) {
    if (views[i].tf_id < 0 || (uint32_t)views[i].tf_id >= vfr_ctx_n_tf(c)) {
      vfr_set_error("view %d: unknown transfer function %d", i, views[i].tf_id);
      return VFR_EINVAL;
    }
    if (o.mode == VFR_MODE_REF && views[i].projection != VFR_PROJ_ORTHO) {
      vfr_set_error("view %d: REF mode renders orthographic views only", i);
      return VFR_EINVAL;
    }
    DevView dv;
    CHECK_RC(vfr_make_view(views[i], lo, hi, &dv));
    pix += (uint64_t)views[i].width * views[i].height;
    tiles += (uint64_t)((views[i].width + 7) / 8) * ((views[i].height + 7) / 8);
    persp += views[i].projection == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
    tf_lo = std::min(tf_lo, (uint32_t)views[i].tf_id);
    tf_hi = std::max(tf_hi, (uint32_t)views[i].tf_id);
  }
  if (pix >= (1ull << 31)) { vfr_set_error("vfr_render_noalloc: more than 2^31 pixels"); return VFR_EINVAL; }
  if ((uint64_t)n_views * d->n_verts >= VFR_MAX_VIEW_VERTS) {
    vfr_set_error("vfr_render_noalloc: %llu view vertices beyond the 32-bit addressing limit",
                  (unsigned long long)n_views * d->n_verts);
    return VFR_EINVAL;
  }
  const uint32_t kmin = o.mode == VFR_MODE_FTB ? (o.max_entries ? o.max_entries : 8u) : 0u;
  Layout L = layout((uint32_t)n_views, pix, d->n_verts, kmin);
  if (L.end > VFR_SCRATCH_BYTES) {
    vfr_set_error("vfr_render_noalloc needs %llu bytes of scratch for %llu pixels and %d views of %u vertices; "
                  "the scratch area holds %llu", (unsigned long long)L.end, (unsigned long long)pix, n_views,
                  d->n_verts, (unsigned long long)VFR_SCRATCH_BYTES);
    return VFR_EBUDGET;
  }
  DevView* dviews = at<DevView>(base, L.views);
  uint32_t* counters = at<uint32_t>(base, L.counters);
  unsigned long long* dstats = at<unsigned long long>(base, L.stats);
  uint32_t* cnt = at<uint32_t>(base, L.cnt);
  uint64_t* rays = at<uint64_t>(base, L.rays);
  float4* rgba = at<float4>(base, L.rgba);
  uint32_t* steps = at<uint32_t>(base, L.steps);
  uint32_t* entry = at<uint32_t>(base, L.entry);
  float4* pverts = at<float4>(base, L.pverts);
  uint64_t* entries = at<uint64_t>(base, L.entries);
  {
    uint64_t p0 = 0, t0v = 0;
    for (int i = 0; i < n_views; ++i) {
      DevView dv;
      CHECK_RC(vfr_make_view(views[i], lo, hi, &dv));
      dv.pix_off = (uint32_t)p0;
      dv.tile_off = (uint32_t)t0v;
      dv.tiles_x = (views[i].width + 7) / 8;
      dv.vert_off = (uint32_t)((uint64_t)i * d->n_verts);
      dv.rec_off = 0;
      dv.ext_off = 0;
      dv.wvert_off = 0;
      dv.n_ext = d->n_ext;
      dv.n_verts = d->n_verts;
      dv.max_edge = d->lmt[0];
      dv.inv_me = 1.0f / d->lmt[0];
      p0 += (uint64_t)views[i].width * views[i].height;
      t0v += (uint64_t)dv.tiles_x * ((views[i].height + 7) / 8);
      CHECK_HIP(hipMemcpyAsync(dviews + i, &dv, sizeof(DevView), hipMemcpyHostToDevice, st));
      CHECK_HIP(hipStreamSynchronize(st)); /* dv is on the stack */
    }
  }
  const uint4* fvid = reinterpret_cast<const uint4*>(d->fvid);
  const float4* wvert = reinterpret_cast<const float4*>(d->vert);
  uint32_t kmax = kmin;
  if (o.mode == VFR_MODE_FTB && o.max_entries == 0) {
    /* the fewest of 8/16/32 slots that hold every pixel's entry faces */
    CHECK_RC(vfr_k_clear_u32(cnt, pix, st));
    CHECK_RC(vfr_k_clear_u32(counters, 1, st));
    CHECK_RC(vfr_k_xform_views(dviews, 0, (uint32_t)n_views, d->n_verts, wvert, pverts, st));
    CHECK_RC(vfr_k_raster_ftb(dviews, 0, (uint32_t)n_views, d->n_ext, fvid, pverts, cnt, nullptr, 0, 0, dstats,
                              nullptr, st));
    CHECK_RC(vfr_k_max_u32(cnt, pix, counters, st));
    uint32_t mx = 0;
    CHECK_HIP(hipMemcpyAsync(&mx, counters, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CHECK_HIP(hipStreamSynchronize(st));
    if (mx > VFR_K_MAX) {
      vfr_set_error("a pixel has %u entry faces, more than the %u entry slots hold; they would be dropped", mx,
                    VFR_K_MAX);
      return VFR_EMESH;
    }
    kmax = mx <= 8 ? 8u : (mx <= 16 ? 16u : VFR_K_MAX);
    L = layout((uint32_t)n_views, pix, d->n_verts, kmax);
    if (L.end > VFR_SCRATCH_BYTES) {
      vfr_set_error("vfr_render_noalloc needs %llu bytes of scratch with %u entry slots per pixel; the scratch "
                    "area holds %llu", (unsigned long long)L.end, kmax, (unsigned long long)VFR_SCRATCH_BYTES);
      return VFR_EBUDGET;
    }
  }

  /* the batched launch set of a plan, on the context's stream */
  CHECK_RC(vfr_k_clear_u32(cnt, pix, st));
  CHECK_RC(vfr_k_clear_u32(counters, (uint32_t)n_views + 4u, st));
  CHECK_RC(vfr_k_clear_u32(reinterpret_cast<uint32_t*>(dstats), 2 * ST_N, st));
  CHECK_RC(vfr_k_xform_views(dviews, 0, (uint32_t)n_views, d->n_verts, wvert, pverts, st));
  if (o.mode == VFR_MODE_REF)
    CHECK_RC(vfr_k_raster_ref(dviews, 0, (uint32_t)n_views, d->n_ext, fvid, pverts, cnt, st));
  else
    CHECK_RC(vfr_k_raster_ftb(dviews, 0, (uint32_t)n_views, d->n_ext, fvid, pverts, cnt, entries, kmax, pix, dstats,
                              nullptr, st));
  CHECK_RC(vfr_k_compact(dviews, (uint32_t)n_views, 0, (uint32_t)tiles, o.mode, cnt, cnt, entries, kmax, pix, rgba,
                         steps, entry, rays, counters + n_views, st));
  TravArgs a;
  memset(&a, 0, sizeof(a));
  a.rec = d->rec;
  a.tv = reinterpret_cast<const uint4*>(d->tv);
  a.vert = pverts;
  a.ext_int = d->ext_int;
  a.fvid = reinterpret_cast<const uint4*>(d->fvid);
  a.tf = vfr_ctx_tf(c);
  a.psi = vfr_ctx_psi(c);
  a.views = dviews;
  a.rays = rays;
  a.n_rays = counters + n_views;
  a.owner = cnt;
  a.cnt = cnt;
  a.entries = entries;
  a.estride = pix;
  a.kmax = kmax;
  a.rgba = rgba;
  a.steps = steps;
  a.entry_out = entry;
  a.stats = dstats;
  a.queue = counters + n_views + 1;
  a.term = o.term_alpha;
  a.refill = o.refill;
  if (tf_hi - tf_lo < VFR_TF_LDS_SLOTS) { a.tf_lds_base = tf_lo; a.tf_lds_n = tf_hi - tf_lo + 1; }
  else a.tf_block = 1; /* each block stages the TFs of its rays' views */
  a.composite = o.composite;
  a.psi_table = (o.flags & VFR_RENDER_PSI_TABLE) ? 1u : 0u;
  a.lds_records = (o.flags & VFR_RENDER_LDS_RECORDS) ? 1u : 0u;
  const uint32_t proj = persp == 0 ? VFR_TRAV_ORTHO : (persp == (uint32_t)n_views ? VFR_TRAV_PERSP : VFR_TRAV_MIXED);
  const int kern = o.mode != VFR_MODE_FTB ? VFR_WALK_THREAD
                   : (o.flags & VFR_RENDER_PERSISTENT) ? VFR_WALK_REFILL
                   : (o.flags & VFR_RENDER_WAVE_QUEUE) ? VFR_WALK_QUEUE
                                                       : VFR_WALK_THREAD;
  const uint64_t records = (o.flags & VFR_RENDER_WIDE) ? VFR_BIG_RECORDS : (uint64_t)d->n_tets;
  CHECK_RC(vfr_k_traverse(o.mode, proj, records, o.occupancy, a, (uint32_t)pix, kern,
                          o.waves_per_cu * (uint32_t)vfr_ctx_cus(c), st));
  unsigned long long s[ST_N];
  uint32_t active = 0;
  CHECK_HIP(hipMemcpyAsync(s, dstats, sizeof(s), hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipMemcpyAsync(&active, counters + n_views, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipStreamSynchronize(st));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (s[ST_OVERFLOW]) {
    vfr_set_error("%llu pixel entries beyond %u per pixel were dropped", s[ST_OVERFLOW], kmax);
    return VFR_EMESH;
  }
  if (s[ST_BROKEN]) {
    vfr_set_error("%llu rays met a neighbour without a back link (non-manifold mesh)", s[ST_BROKEN]);
    return VFR_EMESH;
  }
  if (outs) {
    const uint32_t* h_extf = vfr_mesh_host_extf(m);
    uint64_t p0 = 0;
    for (int i = 0; i < n_views; ++i) {
      const vfr_frame_out& fo = outs[i];
      const size_t n = (size_t)views[i].width * views[i].height;
      if (fo.rgba) {
        CHECK_HIP(hipMemcpy(fo.rgba, rgba + p0, 16 * n, hipMemcpyDeviceToHost));
        if (fo.rgb8)
          for (size_t q = 0; q < n; ++q)
            for (int ch = 0; ch < 3; ++ch) {
              float x = fo.rgba[q * 4 + ch];
              x = x > 1.0f ? 1.0f : x;
              fo.rgb8[q * 3 + ch] = (uint8_t)(int)(255.0f * x);
            }
      }
      if (fo.entry_ext) CHECK_HIP(hipMemcpy(fo.entry_ext, entry + p0, 4 * n, hipMemcpyDeviceToHost));
      if (fo.entry_tet) {
        /* the ids land in the caller's buffer and are mapped in place */
        CHECK_HIP(hipMemcpy(fo.entry_tet, entry + p0, 4 * n, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q)
          fo.entry_tet[q] = fo.entry_tet[q] == VFR_NO_ID ? VFR_NO_ID : h_extf[(size_t)fo.entry_tet[q] * 2];
      }
      if (fo.steps) CHECK_HIP(hipMemcpy(fo.steps, steps + p0, 4 * n, hipMemcpyDeviceToHost));
      p0 += n;
    }
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->rays = pix;
    stats->active_rays = active;
    stats->steps = s[ST_STEPS];
    stats->anomalies = s[ST_ANOM];
    stats->ms_total = ms;
  }
  return VFR_OK;
}
