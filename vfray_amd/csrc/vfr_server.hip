/*
 * vfr_server.hip -- the render server: resident meshes, admission by HBM
 * footprint, continuous batching over a few streams, latency accounting.
 *
 * Slots (VFR_SERVE_SLOTS): each slot owns a stream and device buffers sized
 * to its share of the HBM budget, allocated once.  A dispatcher thread polls
 * the slots' completion events and, whenever a slot is free, admits the
 * queued requests that fit it (oldest first, up to max_batch requests and
 * max_batch_pixels pixels) and enqueues the batched launch set: view upload,
 * clears, per-view vertex transform, visibility/raster, compaction and the
 * traversal.  Batches on different slots overlap on the device.
 *
 * Pipeline (VFR_SERVE_PIPELINE): the per-pixel buffers, the P-space vertices
 * and the views are shared arenas from which each admitted request takes its
 * own ranges (first fit), so admission waits only for memory.  The
 * dispatcher admits the front of the queue as soon as it fits and enqueues
 * its preprocessing on the preprocessing stream; the compaction appends the
 * rays to the open walk list (one ray list, the concatenation of every batch
 * preprocessed since the list opened).  Whenever one of `streams` walk
 * streams is free, the open list is closed and walked in one launch after its
 * batches' preprocessing events; a new list opens.  So requests that arrive
 * while walks run are preprocessed meanwhile and join the next walk, which
 * starts while the previous one drains.  The dispatcher never waits on the
 * device.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
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

#define CHECK_RC(x)                \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != VFR_OK) return rc_; \
  } while (0)

namespace {

using Clock = std::chrono::steady_clock;

inline double secs(Clock::time_point t) {
  return std::chrono::duration<double>(t.time_since_epoch()).count();
}

/* Device bytes a request needs: per pixel the entry count, VFR_K_MAX entry
 * keys, a ray slot, RGBA, steps and entry id; per mesh vertex one P-space
 * vertex.  Every batch gets the most entry slots a plan can have, so that no
 * view a request can ask for needs a counting pass first (the slots beyond a
 * pixel's count cost memory only: the raster, the sort and the walk touch
 * as many as the pixel has). */
/* Per pixel: count, entry slots, ray, RGBA, steps, entry id, and 2 bytes for
 * the tile-ordered compaction (12 bytes per 8x8 tile; a view of W x H pixels
 * has at most W H / 8 + 1 tiles, because (W + 7)(H + 7) / 64 <= W H / 8 + 1
 * whenever (W - 1)(H - 1) >= 0). */
inline uint64_t pixel_bytes() { return 4 + 8ull * VFR_K_MAX + 8 + 16 + 4 + 4 + 2; }

struct SrvMesh {
  uint32_t rec_off, ext_off, wvert_off, n_ext, n_verts, n_tets;
  float max_edge;
  float bb_lo[3], bb_hi[3];
};

struct Req {
  uint64_t ticket;
  int mesh;
  vfr_view view;
  DevView dv;
  uint32_t keep;
  uint64_t pixels;
  uint64_t keep_off = 0;   /* offset of the kept image in the slot's staging */
  double t_arrival = 0, t_admit = 0, t_done = 0;
  int32_t status = 0;
  uint32_t batch = 0;
  uint32_t vflag = 0;      /* VFR_VFLAG_* its view recorded */
  std::vector<float> image;
  /* pipeline: the request's ranges in the shared arenas */
  uint64_t view_idx = 0, pix_off = 0, vert_off = 0;
};

/* First-fit allocator of ranges in [0, cap). */
struct Ranges {
  std::map<uint64_t, uint64_t> free_; /* start -> length */
  void init(uint64_t cap) {
    free_.clear();
    if (cap) free_[0] = cap;
  }
  bool alloc(uint64_t n, uint64_t* off) {
    if (n == 0) { *off = 0; return true; }
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const uint64_t s = it->first + n, l = it->second - n;
        free_.erase(it);
        if (l) free_[s] = l;
        return true;
      }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    if (n == 0) return;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }
};

/* A pipeline walk list: the rays of every batch preprocessed while it was
 * open, walked in one launch. */
struct WalkList {
  enum { FREE = 0, OPEN = 1, WALKING = 2 };
  int state = FREE;
  uint64_t* rays = nullptr;              /* [cap_pix] */
  uint32_t* counters = nullptr;          /* [0] rays, [1] queue head */
  unsigned long long* stats = nullptr;   /* [ST_N] */
  unsigned long long* h_stats = nullptr; /* pinned */
  uint32_t* h_counters = nullptr;        /* pinned */
  uint32_t* h_vflags = nullptr;          /* pinned copy of the view flags */
  float* h_img = nullptr;                /* pinned staging of kept images */
  hipEvent_t done = nullptr;
  std::vector<hipEvent_t> prep_ev;       /* the batches' preprocessing events */
  size_t n_prep = 0;
  std::vector<Req*> reqs;
  uint64_t pix = 0, keep_pix = 0;
  uint32_t n_persp = 0, tf_lo = UINT32_MAX, tf_hi = 0;
  int stream = -1;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint64_t cap_pix = 0, cap_verts = 0;
  uint32_t cap_views = 0;
  uint32_t cap_tiles = 0;       /* cap_pix / 8 + cap_views */
  DevView* views = nullptr;
  DevView* h_views = nullptr;   /* pinned */
  uint32_t* cnt = nullptr;
  uint64_t* entries = nullptr;
  uint64_t* rays = nullptr;
  float4* pverts = nullptr;
  float4* rgba = nullptr;
  uint32_t* steps = nullptr;
  uint32_t* entry = nullptr;
  uint32_t* counters = nullptr; /* [0] rays, [1] queue head */
  uint32_t* block_off = nullptr; /* compaction: per-block ray counts, then offsets */
  uint64_t* tile_mask = nullptr; /* compaction: [cap_tiles] masks, then [cap_tiles] views */
  uint32_t* vflags = nullptr;   /* [cap_views] VFR_VFLAG_* per request */
  uint32_t* h_vflags = nullptr; /* pinned */
  unsigned long long* stats = nullptr;
  unsigned long long* h_stats = nullptr; /* pinned */
  uint32_t* h_counters = nullptr;        /* pinned */
  float* h_img = nullptr;                /* pinned staging for kept images */
  uint64_t cap_keep = 0;                 /* staging capacity in pixels */
  std::vector<Req*> batch;
  bool busy = false;
};

}  // namespace

struct vfr_server {
  vfr_ctx* ctx = nullptr;
  vfr_server_cfg cfg{};
  std::vector<SrvMesh> meshes;
  TetRec* rec = nullptr;
  uint4* tv = nullptr;
  uint64_t n_records = 0;
  uint32_t* ext = nullptr;
  uint4* fvid = nullptr;       /* external faces' vertex ids, same offsets as ext */
  float4* vert = nullptr;
  std::vector<std::unique_ptr<Slot>> slots;
  /* pipeline */
  hipStream_t prep_stream = nullptr;
  std::vector<hipStream_t> walk_streams;
  std::vector<int> walk_busy;
  DevView* p_views = nullptr;     /* [cap_views] */
  DevView* p_h_views = nullptr;   /* pinned staging, same indices */
  uint32_t* p_vflags = nullptr;   /* [cap_views] */
  uint32_t* p_cnt = nullptr;      /* [cap_pix] */
  uint64_t* p_entries = nullptr;  /* [VFR_K_MAX * cap_pix] */
  float4* p_rgba = nullptr;
  uint32_t* p_steps = nullptr;
  uint32_t* p_entry = nullptr;
  float4* p_pverts = nullptr;     /* [cap_verts] */
  uint64_t cap_pix = 0, cap_verts = 0, cap_views = 0, cap_keep = 0;
  Ranges r_pix, r_verts, r_views;
  std::vector<std::unique_ptr<WalkList>> lists;
  std::vector<hipEvent_t> ev_pool;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_ticket = 1;
  std::thread worker;
  bool stop = false;
  int worker_error = VFR_OK;
  /* accounting */
  uint64_t n_batches = 0, n_batched_reqs = 0, steps = 0, anomalies = 0, active_rays = 0;
};

static void slot_free(Slot* s) {
  void* dev[] = {s->views, s->cnt, s->entries, s->rays, s->pverts, s->rgba, s->steps, s->entry,
                 s->counters, s->block_off, s->tile_mask, s->stats, s->vflags};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  void* host[] = {s->h_views, s->h_stats, s->h_counters, s->h_img, s->h_vflags};
  for (void* p : host)
    if (p) (void)hipHostFree(p);
  if (s->done) (void)hipEventDestroy(s->done);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

static int slot_alloc(Slot* s, uint64_t bytes, uint32_t max_batch) {
  /* three quarters for per-pixel buffers, one quarter for P-space vertices */
  s->cap_pix = (bytes * 3 / 4) / pixel_bytes();
  s->cap_verts = std::min<uint64_t>((bytes / 4) / 16, VFR_MAX_VIEW_VERTS - 1);
  s->cap_views = max_batch;
  s->cap_tiles = (uint32_t)std::min<uint64_t>(s->cap_pix / 8 + s->cap_views, UINT32_MAX);
  CHECK_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  CHECK_HIP(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
  CHECK_HIP(hipMalloc((void**)&s->views, sizeof(DevView) * s->cap_views));
  CHECK_HIP(hipMalloc((void**)&s->cnt, 4 * s->cap_pix));
  CHECK_HIP(hipMalloc((void**)&s->entries, 8ull * VFR_K_MAX * s->cap_pix));
  CHECK_HIP(hipMalloc((void**)&s->rays, 8 * s->cap_pix));
  CHECK_HIP(hipMalloc((void**)&s->pverts, 16 * s->cap_verts));
  CHECK_HIP(hipMalloc((void**)&s->rgba, 16 * s->cap_pix));
  CHECK_HIP(hipMalloc((void**)&s->steps, 4 * s->cap_pix));
  CHECK_HIP(hipMalloc((void**)&s->entry, 4 * s->cap_pix));
  CHECK_HIP(hipMalloc((void**)&s->counters, 16));
  CHECK_HIP(hipMalloc((void**)&s->block_off, sizeof(uint32_t) * vfr_compact_blocks(s->cap_tiles)));
  CHECK_HIP(hipMalloc((void**)&s->tile_mask, (sizeof(uint64_t) + sizeof(uint32_t)) * (size_t)s->cap_tiles));
  CHECK_HIP(hipMalloc((void**)&s->stats, sizeof(unsigned long long) * ST_N));
  CHECK_HIP(hipMalloc((void**)&s->vflags, sizeof(uint32_t) * s->cap_views));
  CHECK_HIP(hipHostMalloc((void**)&s->h_views, sizeof(DevView) * s->cap_views, hipHostMallocDefault));
  CHECK_HIP(hipHostMalloc((void**)&s->h_vflags, sizeof(uint32_t) * s->cap_views, hipHostMallocDefault));
  CHECK_HIP(hipHostMalloc((void**)&s->h_stats, sizeof(unsigned long long) * ST_N, hipHostMallocDefault));
  CHECK_HIP(hipHostMalloc((void**)&s->h_counters, 16, hipHostMallocDefault));
  return VFR_OK;
}

/* Enqueues one batch on the slot's stream.  Caller holds no lock. */
static int launch_batch(vfr_server* sv, Slot* s) {
  const uint32_t nv = (uint32_t)s->batch.size();
  uint64_t pix = 0, tiles = 0, verts = 0, keep_pix = 0;
  uint32_t n_persp = 0, tf_lo = UINT32_MAX, tf_hi = 0, max_verts = 0, max_faces = 0;
  for (uint32_t i = 0; i < nv; ++i) {
    Req* r = s->batch[i];
    const SrvMesh& m = sv->meshes[r->mesh];
    DevView d = r->dv;
    n_persp += d.proj == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
    tf_lo = std::min(tf_lo, d.tf);
    tf_hi = std::max(tf_hi, d.tf);
    d.pix_off = (uint32_t)pix;
    d.tile_off = (uint32_t)tiles;
    d.tiles_x = (r->view.width + 7) / 8;
    d.vert_off = (uint32_t)verts;
    s->h_views[i] = d;
    pix += r->pixels;
    tiles += (uint64_t)d.tiles_x * ((r->view.height + 7) / 8);
    verts += m.n_verts;
    max_verts = std::max(max_verts, m.n_verts);
    max_faces = std::max(max_faces, m.n_ext);
    if (r->keep) {
      r->keep_off = keep_pix;
      keep_pix += r->pixels;
    }
  }
  if (tiles > s->cap_tiles) {
    vfr_set_error("server batch: %llu tiles exceed the slot's %u", (unsigned long long)tiles, s->cap_tiles);
    return VFR_EINVAL;
  }
  hipStream_t st = s->stream;
  CHECK_HIP(hipMemcpyAsync(s->views, s->h_views, sizeof(DevView) * nv, hipMemcpyHostToDevice, st));
  CHECK_RC(vfr_k_clear_u32(s->cnt, pix, st));
  CHECK_RC(vfr_k_clear_u32(s->counters, 4, st));
  CHECK_RC(vfr_k_clear_u32((uint32_t*)s->stats, 2 * ST_N, st));
  CHECK_RC(vfr_k_clear_u32(s->vflags, nv, st));
  CHECK_RC(vfr_k_xform_views(s->views, 0, nv, max_verts, sv->vert, s->pverts, st));
  CHECK_RC(vfr_k_raster_ftb(s->views, 0, nv, max_faces, sv->fvid, s->pverts, s->cnt, s->entries, VFR_K_MAX,
                            s->cap_pix, s->stats, s->vflags, st));
  /* rays in tile order (count pass, block scan, emit) */
  CHECK_RC(vfr_k_compact(s->views, nv, 0, (uint32_t)tiles, VFR_MODE_FTB, nullptr, s->cnt, s->entries, VFR_K_MAX,
                         s->cap_pix, s->rgba, s->steps, s->entry, s->rays, s->counters, st, 0, s->block_off,
                         s->tile_mask));
  TravArgs a;
  memset(&a, 0, sizeof(a));
  a.rec = sv->rec;
  a.tv = sv->tv;
  a.vert = s->pverts;
  a.ext_int = sv->ext;
  a.fvid = sv->fvid;
  a.tf = vfr_ctx_tf(sv->ctx);
  a.psi = vfr_ctx_psi(sv->ctx);
  a.views = s->views;
  a.rays = s->rays;
  a.n_rays = s->counters;
  a.cnt = s->cnt;
  a.entries = s->entries;
  a.estride = s->cap_pix;
  a.rgba = s->rgba;
  a.steps = s->steps;
  a.entry_out = s->entry;
  a.stats = s->stats;
  a.queue = s->counters + 1;
  a.term = sv->cfg.term_alpha;
  a.refill = sv->cfg.refill;
  a.kmax = VFR_K_MAX;
  a.vflags = s->vflags;
  a.psi_table = (sv->cfg.flags & VFR_RENDER_PSI_TABLE) ? 1u : 0u;
  a.lds_records = (sv->cfg.flags & VFR_RENDER_LDS_RECORDS) ? 1u : 0u;
  a.composite = sv->cfg.composite;
  if (tf_hi - tf_lo < VFR_TF_LDS_SLOTS) { a.tf_lds_base = tf_lo; a.tf_lds_n = tf_hi - tf_lo + 1; }
  else a.tf_block = 1; /* each block stages the TFs of its rays' views */
  const uint32_t proj = n_persp == 0 ? VFR_TRAV_ORTHO : (n_persp == nv ? VFR_TRAV_PERSP : VFR_TRAV_MIXED);
  const uint64_t records = (sv->cfg.flags & VFR_RENDER_WIDE) ? VFR_BIG_RECORDS : sv->n_records;
  CHECK_RC(vfr_k_traverse(VFR_MODE_FTB, proj, records, sv->cfg.occupancy, a, (uint32_t)pix,
                          (sv->cfg.flags & VFR_RENDER_PERSISTENT)    ? VFR_WALK_REFILL
                          : (sv->cfg.flags & VFR_RENDER_WAVE_QUEUE) ? VFR_WALK_QUEUE
                                                                    : VFR_WALK_THREAD,
                          sv->cfg.waves_per_cu * (uint32_t)vfr_ctx_cus(sv->ctx), st));
  CHECK_HIP(hipMemcpyAsync(s->h_stats, s->stats, sizeof(unsigned long long) * ST_N, hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipMemcpyAsync(s->h_counters, s->counters, 16, hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipMemcpyAsync(s->h_vflags, s->vflags, sizeof(uint32_t) * nv, hipMemcpyDeviceToHost, st));
  if (keep_pix) {
    for (uint32_t i = 0; i < nv; ++i) {
      Req* r = s->batch[i];
      if (!r->keep) continue;
      const DevView& d = s->h_views[i];
      CHECK_HIP(hipMemcpyAsync(s->h_img + 4ull * r->keep_off, s->rgba + d.pix_off, 16ull * r->pixels,
                               hipMemcpyDeviceToHost, st));
    }
  }
  CHECK_HIP(hipEventRecord(s->done, st));
  return VFR_OK;
}

/* Results of a finished slot; caller holds the lock.  A request whose view
 * dropped entry faces (a pixel with more than VFR_K_MAX) or met a
 * non-manifold face ends with VFR_EMESH instead of an image. */
static void finish_batch(vfr_server* sv, Slot* s, double now) {
  for (uint32_t i = 0; i < s->batch.size(); ++i) {
    Req* r = s->batch[i];
    r->t_done = now;
    r->vflag = s->h_vflags[i];
    r->status = r->vflag ? VFR_EMESH : 2;
    if (r->keep && !r->vflag)
      r->image.assign(s->h_img + 4ull * r->keep_off, s->h_img + 4ull * (r->keep_off + r->pixels));
  }
  sv->n_batches++;
  sv->n_batched_reqs += s->batch.size();
  sv->steps += s->h_stats[ST_STEPS];
  sv->anomalies += s->h_stats[ST_ANOM];
  sv->active_rays += s->h_counters[0];
  s->batch.clear();
  s->busy = false;
}

/* ----------------------------------------------------------------- pipeline */

/* Per-pixel bytes of the pipeline's shared arena: entry count, VFR_K_MAX entry
 * keys, RGBA, steps, entry id, and a ray slot in each walk list. */
static uint64_t pipe_pixel_bytes(uint32_t lists) { return 4 + 8ull * VFR_K_MAX + 16 + 4 + 4 + 8ull * lists; }

static int pipe_alloc(vfr_server* sv) {
  const uint32_t nl = sv->cfg.streams + 1; /* one open list and one per walk stream */
  const uint64_t b = sv->cfg.hbm_budget;
  /* three quarters for per-pixel buffers, one quarter for P-space vertices */
  sv->cap_pix = std::min<uint64_t>((b * 3 / 4) / pipe_pixel_bytes(nl), (1ull << 31) - 1);
  sv->cap_verts = std::min<uint64_t>((b / 4) / 16, VFR_MAX_VIEW_VERTS - 1);
  sv->cap_views = std::max<uint64_t>(4ull * sv->cfg.max_batch, 1024);
  sv->cap_keep = std::min<uint64_t>(sv->cap_pix, 1ull << 22); /* 4 Mpixel = 64 MB of RGBA per list */
  sv->r_pix.init(sv->cap_pix);
  sv->r_verts.init(sv->cap_verts);
  sv->r_views.init(sv->cap_views);
  CHECK_HIP(hipStreamCreateWithFlags(&sv->prep_stream, hipStreamNonBlocking));
  sv->walk_streams.assign(sv->cfg.streams, nullptr);
  sv->walk_busy.assign(sv->cfg.streams, 0);
  for (auto& st : sv->walk_streams) CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  CHECK_HIP(hipMalloc((void**)&sv->p_views, sizeof(DevView) * sv->cap_views));
  CHECK_HIP(hipHostMalloc((void**)&sv->p_h_views, sizeof(DevView) * sv->cap_views, hipHostMallocDefault));
  CHECK_HIP(hipMalloc((void**)&sv->p_vflags, sizeof(uint32_t) * sv->cap_views));
  CHECK_HIP(hipMalloc((void**)&sv->p_cnt, 4 * sv->cap_pix));
  CHECK_HIP(hipMalloc((void**)&sv->p_entries, 8ull * VFR_K_MAX * sv->cap_pix));
  CHECK_HIP(hipMalloc((void**)&sv->p_rgba, 16 * sv->cap_pix));
  CHECK_HIP(hipMalloc((void**)&sv->p_steps, 4 * sv->cap_pix));
  CHECK_HIP(hipMalloc((void**)&sv->p_entry, 4 * sv->cap_pix));
  CHECK_HIP(hipMalloc((void**)&sv->p_pverts, 16 * sv->cap_verts));
  for (uint32_t i = 0; i < nl; ++i) {
    sv->lists.emplace_back(new WalkList);
    WalkList* L = sv->lists.back().get();
    CHECK_HIP(hipMalloc((void**)&L->rays, 8 * sv->cap_pix));
    CHECK_HIP(hipMalloc((void**)&L->counters, 16));
    CHECK_HIP(hipMalloc((void**)&L->stats, sizeof(unsigned long long) * ST_N));
    CHECK_HIP(hipHostMalloc((void**)&L->h_stats, sizeof(unsigned long long) * ST_N, hipHostMallocDefault));
    CHECK_HIP(hipHostMalloc((void**)&L->h_counters, 16, hipHostMallocDefault));
    CHECK_HIP(hipHostMalloc((void**)&L->h_vflags, sizeof(uint32_t) * sv->cap_views, hipHostMallocDefault));
    CHECK_HIP(hipHostMalloc((void**)&L->h_img, 16 * sv->cap_keep, hipHostMallocDefault));
    CHECK_HIP(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
  }
  return VFR_OK;
}

static void pipe_free(vfr_server* sv) {
  for (auto& lp : sv->lists) {
    WalkList* L = lp.get();
    void* dev[] = {L->rays, L->counters, L->stats};
    for (void* q : dev)
      if (q) (void)hipFree(q);
    void* host[] = {L->h_stats, L->h_counters, L->h_vflags, L->h_img};
    for (void* q : host)
      if (q) (void)hipHostFree(q);
    if (L->done) (void)hipEventDestroy(L->done);
    for (hipEvent_t e : L->prep_ev) (void)hipEventDestroy(e);
  }
  sv->lists.clear();
  for (hipEvent_t e : sv->ev_pool) (void)hipEventDestroy(e);
  sv->ev_pool.clear();
  void* dev[] = {sv->p_views, sv->p_vflags, sv->p_cnt, sv->p_entries, sv->p_rgba, sv->p_steps, sv->p_entry,
                 sv->p_pverts};
  for (void* q : dev)
    if (q) (void)hipFree(q);
  if (sv->p_h_views) (void)hipHostFree(sv->p_h_views);
  for (hipStream_t st : sv->walk_streams)
    if (st) (void)hipStreamDestroy(st);
  if (sv->prep_stream) (void)hipStreamDestroy(sv->prep_stream);
}

/* A list becomes the open one: its counters and statistics are cleared on the
 * preprocessing stream, before any batch appends to it. */
static int pipe_open(vfr_server* sv, WalkList* L) {
  CHECK_RC(vfr_k_clear_u32(L->counters, 4, sv->prep_stream));
  CHECK_RC(vfr_k_clear_u32((uint32_t*)L->stats, 2 * ST_N, sv->prep_stream));
  L->state = WalkList::OPEN;
  L->n_prep = 0;
  L->reqs.clear();
  L->pix = L->keep_pix = 0;
  L->n_persp = 0;
  L->tf_lo = UINT32_MAX;
  L->tf_hi = 0;
  L->stream = -1;
  return VFR_OK;
}

/* Preprocessing of requests [0, n) of `batch` (their arena ranges taken, views
 * at [view_base, view_base + n)) on the preprocessing stream, appending their
 * rays to the open list L. */
static int pipe_prep(vfr_server* sv, WalkList* L, const std::vector<Req*>& batch, uint64_t view_base) {
  const uint32_t nv = (uint32_t)batch.size();
  uint64_t tiles = 0;
  uint32_t max_verts = 0, max_faces = 0, max_pix = 0;
  for (uint32_t i = 0; i < nv; ++i) {
    Req* r = batch[i];
    const SrvMesh& m = sv->meshes[r->mesh];
    DevView d = r->dv;
    d.pix_off = (uint32_t)r->pix_off;
    d.tile_off = (uint32_t)tiles;
    d.tiles_x = (r->view.width + 7) / 8;
    d.vert_off = (uint32_t)r->vert_off;
    sv->p_h_views[view_base + i] = d;
    tiles += (uint64_t)d.tiles_x * ((r->view.height + 7) / 8);
    max_verts = std::max(max_verts, m.n_verts);
    max_faces = std::max(max_faces, m.n_ext);
    max_pix = std::max<uint32_t>(max_pix, (uint32_t)r->pixels);
  }
  hipStream_t st = sv->prep_stream;
  DevView* dv = sv->p_views + view_base;
  CHECK_HIP(hipMemcpyAsync(dv, sv->p_h_views + view_base, sizeof(DevView) * nv, hipMemcpyHostToDevice, st));
  CHECK_RC(vfr_k_clear_u32(sv->p_vflags + view_base, nv, st));
  CHECK_RC(vfr_k_clear_view_pixels(sv->p_views, (uint32_t)view_base, nv, max_pix, sv->p_cnt, st));
  CHECK_RC(vfr_k_xform_views(sv->p_views, (uint32_t)view_base, nv, max_verts, sv->vert, sv->p_pverts, st));
  CHECK_RC(vfr_k_raster_ftb(sv->p_views, (uint32_t)view_base, nv, max_faces, sv->fvid, sv->p_pverts, sv->p_cnt,
                            sv->p_entries, VFR_K_MAX, sv->cap_pix, L->stats, sv->p_vflags, st));
  CHECK_RC(vfr_k_compact(dv, nv, 0, (uint32_t)tiles, VFR_MODE_FTB, nullptr, sv->p_cnt, sv->p_entries, VFR_K_MAX,
                         sv->cap_pix, sv->p_rgba, sv->p_steps, sv->p_entry, L->rays, L->counters, st,
                         (uint32_t)view_base));
  if (L->n_prep == L->prep_ev.size()) {
    hipEvent_t e = nullptr;
    if (!sv->ev_pool.empty()) {
      e = sv->ev_pool.back();
      sv->ev_pool.pop_back();
    } else {
      CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    L->prep_ev.push_back(e);
  }
  CHECK_HIP(hipEventRecord(L->prep_ev[L->n_prep++], st));
  return VFR_OK;
}

/* Walks list L on walk stream k once its batches are preprocessed, then
 * copies back the statistics, the view flags and the kept images. */
static int pipe_walk(vfr_server* sv, WalkList* L, int k) {
  hipStream_t st = sv->walk_streams[k];
  for (size_t i = 0; i < L->n_prep; ++i) CHECK_HIP(hipStreamWaitEvent(st, L->prep_ev[i], 0));
  TravArgs a;
  memset(&a, 0, sizeof(a));
  a.rec = sv->rec;
  a.tv = sv->tv;
  a.vert = sv->p_pverts;
  a.ext_int = sv->ext;
  a.fvid = sv->fvid;
  a.tf = vfr_ctx_tf(sv->ctx);
  a.psi = vfr_ctx_psi(sv->ctx);
  a.views = sv->p_views;
  a.rays = L->rays;
  a.n_rays = L->counters;
  a.cnt = sv->p_cnt;
  a.entries = sv->p_entries;
  a.estride = sv->cap_pix;
  a.rgba = sv->p_rgba;
  a.steps = sv->p_steps;
  a.entry_out = sv->p_entry;
  a.stats = L->stats;
  a.queue = L->counters + 1;
  a.term = sv->cfg.term_alpha;
  a.refill = sv->cfg.refill;
  a.kmax = VFR_K_MAX;
  a.vflags = sv->p_vflags;
  a.psi_table = (sv->cfg.flags & VFR_RENDER_PSI_TABLE) ? 1u : 0u;
  a.lds_records = (sv->cfg.flags & VFR_RENDER_LDS_RECORDS) ? 1u : 0u;
  a.composite = sv->cfg.composite;
  if (L->tf_hi - L->tf_lo < VFR_TF_LDS_SLOTS) { a.tf_lds_base = L->tf_lo; a.tf_lds_n = L->tf_hi - L->tf_lo + 1; }
  else a.tf_block = 1; /* each block stages the TFs of its rays' views */
  const uint32_t nv = (uint32_t)L->reqs.size();
  const uint32_t proj = L->n_persp == 0 ? VFR_TRAV_ORTHO : (L->n_persp == nv ? VFR_TRAV_PERSP : VFR_TRAV_MIXED);
  const uint64_t records = (sv->cfg.flags & VFR_RENDER_WIDE) ? VFR_BIG_RECORDS : sv->n_records;
  CHECK_RC(vfr_k_traverse(VFR_MODE_FTB, proj, records, sv->cfg.occupancy, a, (uint32_t)L->pix,
                          (sv->cfg.flags & VFR_RENDER_PERSISTENT)    ? VFR_WALK_REFILL
                          : (sv->cfg.flags & VFR_RENDER_WAVE_QUEUE) ? VFR_WALK_QUEUE
                                                                    : VFR_WALK_THREAD,
                          sv->cfg.waves_per_cu * (uint32_t)vfr_ctx_cus(sv->ctx), st));
  CHECK_HIP(hipMemcpyAsync(L->h_stats, L->stats, sizeof(unsigned long long) * ST_N, hipMemcpyDeviceToHost, st));
  CHECK_HIP(hipMemcpyAsync(L->h_counters, L->counters, 16, hipMemcpyDeviceToHost, st));
  /* the flags of every view slot (a few KB; only this list's are read) */
  CHECK_HIP(hipMemcpyAsync(L->h_vflags, sv->p_vflags, sizeof(uint32_t) * sv->cap_views, hipMemcpyDeviceToHost, st));
  for (Req* r : L->reqs) {
    if (r->keep)
      CHECK_HIP(hipMemcpyAsync(L->h_img + 4ull * r->keep_off, sv->p_rgba + r->pix_off, 16ull * r->pixels,
                               hipMemcpyDeviceToHost, st));
  }
  CHECK_HIP(hipEventRecord(L->done, st));
  L->state = WalkList::WALKING;
  L->stream = k;
  return VFR_OK;
}

/* Results of a walked list; caller holds the lock. */
static void pipe_finish(vfr_server* sv, WalkList* L, double now) {
  for (Req* r : L->reqs) {
    r->t_done = now;
    r->vflag = L->h_vflags[r->view_idx];
    r->status = r->vflag ? VFR_EMESH : 2;
    r->batch = (uint32_t)L->reqs.size();
    if (r->keep && !r->vflag)
      r->image.assign(L->h_img + 4ull * r->keep_off, L->h_img + 4ull * (r->keep_off + r->pixels));
    sv->r_pix.release(r->pix_off, r->pixels);
    sv->r_verts.release(r->vert_off, sv->meshes[r->mesh].n_verts);
    sv->r_views.release(r->view_idx, 1);
  }
  sv->n_batches++;
  sv->n_batched_reqs += L->reqs.size();
  sv->steps += L->h_stats[ST_STEPS];
  sv->anomalies += L->h_stats[ST_ANOM];
  sv->active_rays += L->h_counters[0];
  sv->walk_busy[L->stream] = 0;
  L->reqs.clear();
  L->state = WalkList::FREE;
}

static WalkList* pipe_list(vfr_server* sv, int state) {
  for (auto& lp : sv->lists)
    if (lp->state == state) return lp.get();
  return nullptr;
}

static void pipe_fail(vfr_server* sv, std::vector<Req*>& reqs, int rc) {
  sv->worker_error = rc;
  for (Req* r : reqs) r->status = rc;
  sv->cv_done.notify_all();
}

static void dispatcher_pipeline(vfr_server* sv) {
  (void)hipSetDevice(vfr_ctx_device(sv->ctx));
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    bool any_busy = false;
    const double now = secs(Clock::now());
    /* completed walks */
    for (auto& lp : sv->lists) {
      WalkList* L = lp.get();
      if (L->state != WalkList::WALKING) continue;
      hipError_t q = hipEventQuery(L->done);
      if (q == hipSuccess) {
        pipe_finish(sv, L, now);
        sv->cv_done.notify_all();
      } else if (q != hipErrorNotReady) {
        vfr_set_error("walk failed: %s", hipGetErrorString(q));
        pipe_fail(sv, L->reqs, VFR_EHIP);
        sv->walk_busy[L->stream] = 0;
        L->reqs.clear();
        L->state = WalkList::FREE;
      } else {
        any_busy = true;
      }
    }
    WalkList* open = pipe_list(sv, WalkList::OPEN);
    if (!open && !sv->queue.empty()) {
      WalkList* f = pipe_list(sv, WalkList::FREE);
      if (f) {
        int rc = pipe_open(sv, f);
        if (rc != VFR_OK) {
          std::vector<Req*> q(sv->queue.begin(), sv->queue.end());
          sv->queue.clear();
          pipe_fail(sv, q, rc);
        } else {
          open = f;
        }
      }
    }
    /* admission: the front of the queue, as long as its ranges fit */
    if (open && !sv->queue.empty()) {
      std::vector<Req*> batch;
      while (!sv->queue.empty() && batch.size() < sv->cfg.max_batch) {
        Req* r = sv->queue.front();
        const SrvMesh& m = sv->meshes[r->mesh];
        const uint64_t kp = r->keep ? r->pixels : 0;
        if (open->keep_pix + kp > sv->cap_keep) break;
        if (sv->cfg.max_batch_pixels && !open->reqs.empty() && open->pix + r->pixels > sv->cfg.max_batch_pixels) break;
        uint64_t po = 0, vo = 0;
        if (!sv->r_pix.alloc(r->pixels, &po)) break;
        if (!sv->r_verts.alloc(m.n_verts, &vo)) { sv->r_pix.release(po, r->pixels); break; }
        r->pix_off = po;
        r->vert_off = vo;
        r->keep_off = open->keep_pix;
        open->keep_pix += kp;
        open->pix += r->pixels;
        open->reqs.push_back(r);
        batch.push_back(r);
        sv->queue.pop_front();
      }
      uint64_t vb = 0;
      if (!batch.empty() && !sv->r_views.alloc(batch.size(), &vb)) {
        /* no contiguous view slots: give the batch back to the queue */
        for (size_t i = batch.size(); i-- > 0;) {
          Req* r = batch[i];
          sv->r_pix.release(r->pix_off, r->pixels);
          sv->r_verts.release(r->vert_off, sv->meshes[r->mesh].n_verts);
          open->keep_pix -= r->keep ? r->pixels : 0;
          open->pix -= r->pixels;
          open->reqs.pop_back();
          sv->queue.push_front(r);
        }
        batch.clear();
      }
      if (!batch.empty()) {
        const double ta = secs(Clock::now());
        for (size_t i = 0; i < batch.size(); ++i) {
          Req* r = batch[i];
          r->view_idx = vb + i;
          r->t_admit = ta;
          r->status = 1;
          open->n_persp += r->dv.proj == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
          open->tf_lo = std::min(open->tf_lo, r->dv.tf);
          open->tf_hi = std::max(open->tf_hi, r->dv.tf);
        }
        lk.unlock();
        int rc = pipe_prep(sv, open, batch, vb);
        lk.lock();
        if (rc != VFR_OK) {
          /* the failed batch leaves the open list (it was appended last), so
           * the list's walk does not report these requests as rendered */
          pipe_fail(sv, batch, rc);
          for (Req* r : batch) {
            sv->r_pix.release(r->pix_off, r->pixels);
            sv->r_verts.release(r->vert_off, sv->meshes[r->mesh].n_verts);
            open->keep_pix -= r->keep ? r->pixels : 0;
            open->pix -= r->pixels;
            open->n_persp -= r->dv.proj == VFR_PROJ_PERSPECTIVE ? 1u : 0u;
          }
          sv->r_views.release(vb, batch.size());
          open->reqs.resize(open->reqs.size() - batch.size());
        }
      }
    }
    /* walk the open list on a free walk stream */
    open = pipe_list(sv, WalkList::OPEN);
    if (open && !open->reqs.empty()) {
      int k = -1;
      for (size_t i = 0; i < sv->walk_busy.size(); ++i)
        if (!sv->walk_busy[i]) { k = (int)i; break; }
      if (k >= 0) {
        sv->walk_busy[k] = 1;
        lk.unlock();
        int rc = pipe_walk(sv, open, k);
        lk.lock();
        if (rc != VFR_OK) {
          pipe_fail(sv, open->reqs, rc);
          sv->walk_busy[k] = 0;
          open->reqs.clear();
          open->state = WalkList::FREE;
        }
        any_busy = true;
      }
    }
    open = pipe_list(sv, WalkList::OPEN);
    const bool pending = !sv->queue.empty() || (open && !open->reqs.empty());
    if (sv->stop && !any_busy && !pending) break;
    if (!any_busy && !pending) {
      sv->cv_work.wait(lk, [sv] { return sv->stop || !sv->queue.empty(); });
    } else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    }
  }
}

static void dispatcher(vfr_server* sv) {
  if (sv->cfg.schedule == VFR_SERVE_PIPELINE) {
    dispatcher_pipeline(sv);
    return;
  }
  (void)hipSetDevice(vfr_ctx_device(sv->ctx));
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    bool any_busy = false;
    const double now = secs(Clock::now());
    for (auto& sp : sv->slots) {
      Slot* s = sp.get();
      if (!s->busy) continue;
      hipError_t q = hipEventQuery(s->done);
      if (q == hipSuccess) {
        finish_batch(sv, s, now);
        sv->cv_done.notify_all();
      } else if (q != hipErrorNotReady) {
        vfr_set_error("batch failed: %s", hipGetErrorString(q));
        sv->worker_error = VFR_EHIP;
        for (Req* r : s->batch) r->status = VFR_EHIP;
        s->batch.clear();
        s->busy = false;
        sv->cv_done.notify_all();
      } else {
        any_busy = true;
      }
    }
    /* admission: oldest requests first into each free slot */
    for (auto& sp : sv->slots) {
      Slot* s = sp.get();
      if (s->busy || sv->queue.empty()) continue;
      uint64_t pix = 0, verts = 0, keep = 0;
      while (!sv->queue.empty() && s->batch.size() < s->cap_views) {
        Req* r = sv->queue.front();
        const SrvMesh& m = sv->meshes[r->mesh];
        const uint64_t kp = r->keep ? r->pixels : 0;
        if (pix + r->pixels > s->cap_pix || verts + m.n_verts > s->cap_verts || keep + kp > s->cap_keep) break;
        if (!s->batch.empty() && sv->cfg.max_batch_pixels && pix + r->pixels > sv->cfg.max_batch_pixels) break;
        pix += r->pixels;
        verts += m.n_verts;
        keep += kp;
        r->t_admit = secs(Clock::now());
        r->status = 1;
        s->batch.push_back(r);
        sv->queue.pop_front();
      }
      if (s->batch.empty()) continue;
      for (Req* r : s->batch) r->batch = (uint32_t)s->batch.size();
      s->busy = true;
      any_busy = true;
      lk.unlock();
      int rc = launch_batch(sv, s);
      lk.lock();
      if (rc != VFR_OK) {
        sv->worker_error = rc;
        for (Req* r : s->batch) r->status = rc;
        s->batch.clear();
        s->busy = false;
        sv->cv_done.notify_all();
      }
    }
    if (sv->stop && !any_busy && sv->queue.empty()) break;
    if (!any_busy && sv->queue.empty()) {
      sv->cv_work.wait(lk, [sv] { return sv->stop || !sv->queue.empty(); });
    } else {
      /* device work in flight: poll its events at a fine grain */
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    }
  }
}

extern "C" void vfr_server_destroy(vfr_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipSetDevice(vfr_ctx_device(sv->ctx));
  for (auto& sp : sv->slots) {
    if (sp->stream) (void)hipStreamSynchronize(sp->stream);
    slot_free(sp.get());
  }
  if (sv->prep_stream) (void)hipStreamSynchronize(sv->prep_stream);
  for (hipStream_t st : sv->walk_streams)
    if (st) (void)hipStreamSynchronize(st);
  pipe_free(sv);
  if (sv->rec) (void)hipFree(sv->rec);
  if (sv->tv) (void)hipFree(sv->tv);
  if (sv->ext) (void)hipFree(sv->ext);
  if (sv->fvid) (void)hipFree(sv->fvid);
  if (sv->vert) (void)hipFree(sv->vert);
  delete sv;
}

extern "C" int vfr_server_create(vfr_ctx* ctx, const vfr_server_cfg* cfg, const vfr_mesh* const* meshes,
                                 int n_meshes, vfr_server** out) {
  if (!ctx || !meshes || n_meshes < 1 || !out) { vfr_set_error("vfr_server_create: bad argument"); return VFR_EINVAL; }
  *out = nullptr;
  vfr_server_cfg c{};
  if (cfg) c = *cfg;
  if (c.hbm_budget == 0) c.hbm_budget = 16ull << 30;
  if (c.streams == 0) c.streams = 2;
  if (c.max_batch == 0) c.max_batch = 256;
  if (!(c.term_alpha > 0.0f)) c.term_alpha = 0.99f;
  if (c.waves_per_cu == 0) {
    uint32_t wps = 0;
    CHECK_RC(vfr_k_traverse_waves_per_simd(c.occupancy, &wps));
    c.waves_per_cu = std::min(64u, 8u * wps);
  }
  if (c.refill == 0) c.refill = 32;
  if (c.streams > 16 || c.max_batch > 4096 || c.term_alpha > 1.0f || c.waves_per_cu > 64 || c.refill > 64 ||
      c.composite > VFR_COMPOSITE_LINEAR_OPACITY || c.schedule > VFR_SERVE_PIPELINE) {
    vfr_set_error("vfr_server_create: configuration out of range");
    return VFR_EINVAL;
  }
  uint64_t tets = 0, exts = 0, verts = 0;
  for (int i = 0; i < n_meshes; ++i) {
    if (!meshes[i] || vfr_mesh_ctx(meshes[i]) != ctx) { vfr_set_error("mesh %d is null or from another context", i); return VFR_EINVAL; }
    const DevMesh* d = vfr_mesh_dev(meshes[i]);
    tets += d->n_tets;
    exts += d->n_ext;
    verts += d->n_verts;
  }
  if (tets >= VFR_MAX_TETS || exts >= (1ull << 31) || verts >= (1ull << 31)) {
    vfr_set_error("resident meshes exceed %u tets, or 2^31 faces or vertices", VFR_MAX_TETS - 1u);
    return VFR_EINVAL;
  }
  vfr_server* sv = new (std::nothrow) vfr_server;
  if (!sv) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  sv->ctx = ctx;
  sv->cfg = c;
  CHECK_HIP(hipSetDevice(vfr_ctx_device(ctx)));
  int rc = VFR_OK;
  if (hipMalloc((void**)&sv->rec, sizeof(TetRec) * tets) != hipSuccess ||
      hipMalloc((void**)&sv->tv, sizeof(uint4) * tets) != hipSuccess ||
      hipMalloc((void**)&sv->ext, 8 * (exts + 1)) != hipSuccess ||
      hipMalloc((void**)&sv->fvid, 16 * (exts + 1)) != hipSuccess ||
      hipMalloc((void**)&sv->vert, 16 * verts) != hipSuccess) {
    vfr_set_error("resident mesh arena allocation failed");
    rc = VFR_ENOMEM;
  }
  uint64_t ro = 0, eo = 0, vo = 0;
  for (int i = 0; rc == VFR_OK && i < n_meshes; ++i) {
    const DevMesh* d = vfr_mesh_dev(meshes[i]);
    SrvMesh m;
    m.rec_off = (uint32_t)ro; m.ext_off = (uint32_t)eo; m.wvert_off = (uint32_t)vo;
    m.n_ext = d->n_ext; m.n_verts = d->n_verts; m.n_tets = d->n_tets;
    m.max_edge = d->lmt[0];
    vfr_mesh_bbox(meshes[i], m.bb_lo, m.bb_hi);
    if (hipMemcpy(sv->rec + ro, d->rec, sizeof(TetRec) * d->n_tets, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(sv->tv + ro, d->tv, sizeof(uint4) * d->n_tets, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(sv->ext + 2 * eo, d->ext_int, 8ull * d->n_ext, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(sv->fvid + eo, d->fvid, 16ull * d->n_ext, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(sv->vert + vo, d->vert, 16ull * d->n_verts, hipMemcpyDeviceToDevice) != hipSuccess) {
      vfr_set_error("copy of mesh %d into the resident arena failed", i);
      rc = VFR_EHIP;
    }
    /* neighbour words index the whole arena */
    if (rc == VFR_OK) rc = vfr_k_rebase_links(sv->rec + ro, d->n_tets, (uint32_t)ro, nullptr);
    if (rc == VFR_OK && hipDeviceSynchronize() != hipSuccess) {
      vfr_set_error("rebasing mesh %d in the resident arena failed", i);
      rc = VFR_EHIP;
    }
    sv->meshes.push_back(m);
    ro += d->n_tets; eo += d->n_ext; vo += d->n_verts;
    sv->n_records = ro;
  }
  if (rc == VFR_OK && c.schedule == VFR_SERVE_PIPELINE) {
    rc = pipe_alloc(sv);
    if (rc == VFR_ENOMEM || rc == VFR_EHIP) vfr_set_error("pipeline arena allocation failed: %s", vfr_last_error());
  }
  for (uint32_t i = 0; rc == VFR_OK && c.schedule == VFR_SERVE_SLOTS && i < c.streams; ++i) {
    sv->slots.emplace_back(new Slot);
    rc = slot_alloc(sv->slots.back().get(), c.hbm_budget / c.streams, c.max_batch);
    Slot* sl = sv->slots.back().get();
    sl->cap_keep = std::min<uint64_t>(sl->cap_pix, 1ull << 22); /* 4 Mpixel = 64 MB of RGBA */
    if (rc == VFR_OK && hipHostMalloc((void**)&sl->h_img, 16 * sl->cap_keep, hipHostMallocDefault) != hipSuccess) {
      vfr_set_error("pinned staging allocation failed");
      rc = VFR_ENOMEM;
    }
  }
  if (rc != VFR_OK) { vfr_server_destroy(sv); return rc; }
  sv->worker = std::thread(dispatcher, sv);
  *out = sv;
  return VFR_OK;
}

static int submit_at(vfr_server* sv, const vfr_request* rq, double t_arrival, uint64_t* ticket) {
  if (!sv || !rq || !ticket) { vfr_set_error("vfr_server_submit: bad argument"); return VFR_EINVAL; }
  if (rq->mesh < 0 || (size_t)rq->mesh >= sv->meshes.size()) { vfr_set_error("unknown mesh %d", rq->mesh); return VFR_EINVAL; }
  if (rq->view.tf_id < 0 || (uint32_t)rq->view.tf_id >= vfr_ctx_n_tf(sv->ctx)) {
    vfr_set_error("unknown transfer function %d", rq->view.tf_id);
    return VFR_EINVAL;
  }
  const SrvMesh& m = sv->meshes[rq->mesh];
  std::unique_ptr<Req> r(new (std::nothrow) Req);
  if (!r) { vfr_set_error("out of memory"); return VFR_ENOMEM; }
  CHECK_RC(vfr_make_view(rq->view, m.bb_lo, m.bb_hi, &r->dv));
  r->dv.rec_off = m.rec_off;
  r->dv.ext_off = m.ext_off;
  r->dv.wvert_off = m.wvert_off;
  r->dv.n_ext = m.n_ext;
  r->dv.n_verts = m.n_verts;
  r->dv.max_edge = m.max_edge;
  r->dv.inv_me = 1.0f / m.max_edge;
  r->mesh = rq->mesh;
  r->view = rq->view;
  r->keep = rq->keep_image;
  r->pixels = (uint64_t)rq->view.width * rq->view.height;
  uint64_t cap_pix, cap_verts, cap_keep;
  if (sv->cfg.schedule == VFR_SERVE_PIPELINE) {
    cap_pix = sv->cap_pix; cap_verts = sv->cap_verts; cap_keep = sv->cap_keep;
  } else {
    const Slot* s0 = sv->slots[0].get();
    cap_pix = s0->cap_pix; cap_verts = s0->cap_verts; cap_keep = s0->cap_keep;
  }
  if (r->pixels > cap_pix || m.n_verts > cap_verts || (r->keep && r->pixels > cap_keep)) {
    vfr_set_error("request of %llu pixels on a %u-vertex mesh exceeds the server's share of the HBM budget",
                  (unsigned long long)r->pixels, m.n_verts);
    return VFR_EBUDGET;
  }
  r->t_arrival = t_arrival;
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->worker_error != VFR_OK) { vfr_set_error("server stopped after an error"); return sv->worker_error; }
  r->ticket = sv->next_ticket++;
  *ticket = r->ticket;
  sv->queue.push_back(r.get());
  sv->reqs.emplace(r->ticket, std::move(r));
  sv->cv_work.notify_one();
  return VFR_OK;
}

extern "C" int vfr_server_submit(vfr_server* sv, const vfr_request* rq, uint64_t* ticket) {
  return submit_at(sv, rq, secs(Clock::now()), ticket);
}

static void fill_result(const Req* r, vfr_result* out) {
  out->status = r->status;
  out->width = r->view.width;
  out->height = r->view.height;
  out->batch_size = r->batch;
  out->latency_ms = r->status == 2 ? (r->t_done - r->t_arrival) * 1e3 : 0.0;
  out->queue_ms = r->status >= 1 ? (r->t_admit - r->t_arrival) * 1e3 : 0.0;
}

extern "C" int vfr_server_poll(vfr_server* sv, uint64_t ticket, vfr_result* out) {
  if (!sv || !out) { vfr_set_error("vfr_server_poll: bad argument"); return VFR_EINVAL; }
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) { vfr_set_error("unknown ticket %llu", (unsigned long long)ticket); return VFR_EINVAL; }
  fill_result(it->second.get(), out);
  return VFR_OK;
}

extern "C" int vfr_server_wait(vfr_server* sv, uint64_t ticket, vfr_result* out, float* rgba_out) {
  if (!sv) { vfr_set_error("vfr_server_wait: bad argument"); return VFR_EINVAL; }
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) { vfr_set_error("unknown ticket %llu", (unsigned long long)ticket); return VFR_EINVAL; }
  Req* r = it->second.get();
  sv->cv_done.wait(lk, [r] { return r->status == 2 || r->status < 0; });
  if (out) fill_result(r, out);
  int rc = r->status < 0 ? r->status : VFR_OK;
  if (rc == VFR_EMESH && r->vflag)
    vfr_set_error("request %llu: %s", (unsigned long long)ticket,
                  (r->vflag & VFR_VFLAG_OVERFLOW)
                      ? "a pixel has more entry faces than the 32 entry slots hold (entries would be dropped)"
                      : "a ray met a neighbour without a back link (non-manifold mesh)");
  if (rc == VFR_OK && rgba_out) {
    if (!r->keep) { vfr_set_error("request was submitted without keep_image"); rc = VFR_ESTATE; }
    else memcpy(rgba_out, r->image.data(), sizeof(float) * r->image.size());
  }
  sv->reqs.erase(it);
  return rc;
}

static double pctl(std::vector<double>& v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  double pos = q * (v.size() - 1);
  size_t lo = (size_t)pos, hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - lo) * (v[hi] - v[lo]);
}

extern "C" int vfr_server_replay(vfr_server* sv, const vfr_request* reqs, const double* offsets, int n,
                                 vfr_server_report* rep) {
  if (!sv || !reqs || !offsets || n < 1 || !rep) { vfr_set_error("vfr_server_replay: bad argument"); return VFR_EINVAL; }
  for (int i = 1; i < n; ++i)
    if (offsets[i] < offsets[i - 1]) { vfr_set_error("arrival offsets must be non-decreasing"); return VFR_EINVAL; }
  uint64_t b0, s0, a0, an0, ar0;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    b0 = sv->n_batches; s0 = sv->steps; a0 = sv->anomalies; an0 = sv->n_batched_reqs; ar0 = sv->active_rays;
  }
  std::vector<uint64_t> tickets(n);
  const auto t0 = Clock::now();
  const double t0s = secs(t0);
  int rc = VFR_OK;
  int submitted = 0;
  for (int i = 0; i < n && rc == VFR_OK; ++i) {
    std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(offsets[i])));
    rc = submit_at(sv, &reqs[i], t0s + offsets[i], &tickets[i]);
    if (rc == VFR_OK) submitted = i + 1;
  }
  if (rc != VFR_OK) {
    /* reap what was submitted before the failure, keeping its message */
    std::string msg = vfr_last_error();
    for (int i = 0; i < submitted; ++i) (void)vfr_server_wait(sv, tickets[i], nullptr, nullptr);
    vfr_set_error("%s", msg.c_str());
    return rc;
  }
  std::vector<double> lat(n);
  double last = 0;
  uint64_t rays = 0;
  int first_err = VFR_OK;
  std::string err_msg;
  for (int i = 0; i < n; ++i) {
    vfr_result r;
    double t_done = 0;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      auto it = sv->reqs.find(tickets[i]);
      if (it == sv->reqs.end()) {
        /* taken by a concurrent vfr_server_wait on the same ticket */
        if (first_err == VFR_OK) {
          first_err = VFR_EINVAL;
          err_msg = "ticket " + std::to_string(tickets[i]) + " was waited for by another caller";
        }
        continue;
      }
      Req* q = it->second.get();
      sv->cv_done.wait(lk, [q] { return q->status == 2 || q->status < 0; });
      t_done = q->t_done;
    }
    int rc2 = vfr_server_wait(sv, tickets[i], &r, nullptr);
    if (rc2 != VFR_OK) {
      if (first_err == VFR_OK) { first_err = rc2; err_msg = vfr_last_error(); }
      continue;
    }
    lat[i] = r.latency_ms;
    last = std::max(last, t_done);
    rays += (uint64_t)r.width * r.height;
  }
  memset(rep, 0, sizeof(*rep));
  rep->requests = (uint64_t)n;
  rep->rays = rays;
  rep->makespan_s = last - (t0s + offsets[0]);
  rep->mrays_per_s = rays / rep->makespan_s / 1e6;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    rep->batches = sv->n_batches - b0;
    rep->steps = sv->steps - s0;
    rep->anomalies = sv->anomalies - a0;
    rep->active_rays = sv->active_rays - ar0;
    rep->mean_batch = rep->batches ? (double)(sv->n_batched_reqs - an0) / rep->batches : 0.0;
  }
  if (first_err != VFR_OK) {
    vfr_set_error("%s", err_msg.c_str());
    return first_err;
  }
  rep->latency_p50_ms = pctl(lat, 0.50);
  rep->latency_p90_ms = pctl(lat, 0.90);
  rep->latency_p99_ms = pctl(lat, 0.99);
  rep->latency_max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
  return VFR_OK;
}
