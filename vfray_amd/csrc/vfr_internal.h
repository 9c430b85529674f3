/*
 * vfr_internal.h -- structures shared by the host API and the HIP kernels.
 */
#ifndef VFR_INTERNAL_H
#define VFR_INTERNAL_H

#include <cstddef>
#include <cstdint>

void vfr_set_error(const char* fmt, ...);
double vfr_psi_integral(double tf, double tb);

/* Packed neighbour word of a tet face: (next_tet << 2) | entry_face_in_next,
 * or one of these codes. */
#define VFR_NB_EXT 0xFFFFFFFFu    /* external face (the .con file's "itself") */
#define VFR_NB_BROKEN 0xFFFFFFFEu /* neighbour whose faces do not link back   */

/* A walk that visits more tets than this is stopped as an anomaly (in REF
 * mode a ray through an edge can circle it forever, as in the reference).  A
 * straight ray through a 400^3 grid crosses about 2,100 tets. */
#define VFR_MAX_STEPS 8192u

/* Neighbour words hold next << 2, so a mesh, or a server's resident meshes
 * together, stay below VFR_MAX_TETS tets; the walk addresses P-space
 * vertices (16 B each) with 32-bit byte offsets, so the vertices of all
 * views of one launch stay below VFR_MAX_VIEW_VERTS. */
#define VFR_MAX_TETS (1u << 29)
/* Record arenas below 4 GiB (32 B per tet) are addressed with 32-bit byte
 * offsets by the fastest traversal kernels. */
#define VFR_BIG_RECORDS (1ull << 27)
#define VFR_MAX_VIEW_VERTS (1ull << 28)

/* View chunks of a pipelined batched render. */
#define VFR_PIPE_CHUNKS 4u

/* Transfer-function slots a traversal launch can stage in LDS (4 KB each). */
#define VFR_TF_LDS_SLOTS 4u

/* Entry-list slots per pixel in front-to-back mode. */
#define VFR_K 8u
/* Most slots a plan's entry lists can have (plans pick 8, 16 or 32). */
#define VFR_K_MAX 32u

#define VFR_NO_ID 0xFFFFFFFFu

/* 32-byte face record, one per tet, 32-byte aligned: per face f (the
 * vertices of slots f, f+1, f+2 of the tet's vertex ids, kept beside in a
 * uint4 array) the packed neighbour word and the face's vertex word: the
 * neighbour's vertex off that face (the tet's own opposite vertex when there
 * is no linked neighbour) in the low VFR_REC_VERT_BITS bits and above them
 * the index into VFR_SLOT_PERMS of the slot map across the face: this tet's
 * slot of each vertex of the face goes to that vertex's slot in the
 * neighbour, and slot f + 3 (the vertex off the face) to the neighbour's own
 * vertex off the face.  A front-to-back step reads only face records (two
 * loads from one line; four tets per 128-byte line): it carries the slots of
 * its face vertices, one per byte, instead of their ids, and moves them into
 * the neighbour's numbering with one byte permute. */
struct alignas(32) TetRec {
  uint32_t f[8];   /* face k: f[2k] neighbour word, f[2k+1] vertex word */
};
#define VFR_REC_VERT_BITS 26u
#define VFR_REC_VERT_MASK ((1u << VFR_REC_VERT_BITS) - 1u)
/* A mesh's vertex ids fit the vertex word. */
#define VFR_MAX_VERTS (1u << VFR_REC_VERT_BITS)
/* The 24 permutations of four slots, lexicographic; byte i = the image of
 * slot i. */
#define VFR_N_SLOT_PERMS 24u
static constexpr uint32_t VFR_SLOT_PERMS[VFR_N_SLOT_PERMS] = {
    0x03020100u, 0x02030100u, 0x03010200u, 0x01030200u, 0x02010300u, 0x01020300u,
    0x03020001u, 0x02030001u, 0x03000201u, 0x00030201u, 0x02000301u, 0x00020301u,
    0x03010002u, 0x01030002u, 0x03000102u, 0x00030102u, 0x01000302u, 0x00010302u,
    0x02010003u, 0x01020003u, 0x02000103u, 0x00020103u, 0x01000203u, 0x00010203u};
static_assert(sizeof(TetRec) == 32, "TetRec is 32 bytes");

/* One view of a render batch, with the offsets of its mesh in the device
 * arenas (a plan's single mesh, or a server's resident meshes). */
struct alignas(16) DevView {
  float M[12];        /* row-major 3x4 view transform                       */
  float k;            /* ortho: world length per unit Z; persp: tan(fov_y/2) */
  uint32_t W, H;
  uint32_t proj;      /* VFR_PROJ_*                                          */
  uint32_t tf;        /* transfer-function slot                              */
  uint32_t pix_off;   /* first pixel of this view in the per-pixel arrays    */
  uint32_t tile_off;  /* first 8x8 tile of this view                         */
  uint32_t tiles_x;   /* 8x8 tiles per image row                             */
  uint32_t vert_off;  /* first P-space vertex of this view (prefix of n_verts) */
  uint32_t rec_off;   /* mesh: first tet record in the record arena          */
  uint32_t ext_off;   /* mesh: first external face in the face arena         */
  uint32_t wvert_off; /* mesh: first world vertex in the vertex arena        */
  uint32_t n_ext;     /* mesh: external faces                                */
  uint32_t n_verts;   /* mesh: vertices                                      */
  float max_edge;     /* mesh: .lmt max edge length                          */
  float inv_me;       /* 1 / max_edge                                        */
  float kx;           /* persp: tan(fov_y/2) * W / H (horizontal half-extent) */
  uint32_t pad_[3];
};
static_assert(sizeof(DevView) % 16 == 0, "DevView alignment");

struct DevMesh {
  uint32_t n_verts = 0, n_tets = 0, n_ext = 0;
  float lmt[3] = {0, 0, 0};
  TetRec* rec = nullptr;       /* [n_tets] face records, internal order     */
  uint32_t* tv = nullptr;      /* [n_tets*4] vertex ids (internal), same order */
  float* vert = nullptr;       /* [n_verts*4] internal order (x,y,z,s)      */
  uint32_t* ext_int = nullptr; /* [n_ext*2] (internal tet, face)            */
  uint32_t* fvid = nullptr;    /* [n_ext*4] internal vertex ids of each external
                                  face's slots f, f+1, f+2 and the one off it */
  uint32_t* extf = nullptr;    /* [n_ext*2] (caller tet, face), .extf order */
  uint32_t* perm = nullptr;    /* [n_tets] internal -> caller tet id        */
  uint32_t* vperm = nullptr;   /* [n_verts] internal -> caller vertex id    */
  uint32_t* con = nullptr;     /* [n_tets*4] caller ids, .con body          */
  uint64_t bytes = 0;
  double build_ms = 0;
};

/* normalizeVertices' centre and scales (vfRayPreComp.cc:113-130) from the
 * componentwise min/max of (x, y, z, s): centre = (min + max) / 2.0 in
 * double, sc = 1 / the largest half extent, ss = 1 / (max s - min s). */
void vfr_normalize_params(const float mn[4], const float mx[4], float center[3], float* sc, float* ss);

/* Statistics counters a render accumulates on the device (ST_BROKEN: rays
 * that met a neighbour whose faces do not link back). */
enum { ST_STEPS = 0, ST_ANOM = 1, ST_OVERFLOW = 2, ST_RAYS = 3, ST_BROKEN = 4, ST_N = 8 };

/* Per-view error flags a launch can record (TravArgs.vflags and the FTB
 * raster's vflags, one word per view of the launch). */
#define VFR_VFLAG_OVERFLOW 1u /* a pixel had more entry faces than its slots */
#define VFR_VFLAG_BROKEN 2u   /* a ray met a neighbour without a back link   */

#ifdef __HIPCC__
struct TravArgs {
  const TetRec* rec;       /* face-record arena; a view's mesh starts at rec_off */
  const uint4* tv;         /* tet vertex ids, same indexing as rec            */
  const float4* vert;      /* P-space vertices of all views (+ view vert_off) */
  const uint32_t* ext_int; /* (internal tet, face) arena; + ext_off           */
  const uint4* fvid;       /* per external face: its tet's vertex ids in slots
                              f, f+1, f+2, f+3 (the face, then the vertex off
                              it); same indexing as ext_int                  */
  const float4* tf;        /* [slots * 256] */
  const float* psi;        /* [512 * 512]   */
  const DevView* views;
  const uint64_t* rays;    /* (view << 32) | pixel */
  const uint32_t* n_rays;  /* device count of rays */
  const uint32_t* owner;   /* REF: ext index + 1, 0 = none */
  const uint32_t* cnt;     /* FTB: entries per pixel */
  const uint64_t* entries; /* FTB: keys (z << 32 | ext), slot-major: slot k of
                              pixel g at entries[k * estride + g] */
  float4* rgba;
  uint32_t* steps;
  uint32_t* entry_out;
  unsigned long long* stats;
  uint32_t* queue;         /* persistent kernel work-queue head */
  float term;
  float term_tr;           /* front to back: the largest T with 1 - T >= term in
                              float arithmetic (set by vfr_k_traverse)         */
  uint32_t refill;         /* persistent: refill when this many lanes idle */
  uint32_t tf_lds_base;    /* one-ray-per-thread FTB: TF slots [base, base+n)   */
  uint32_t tf_lds_n;       /* staged in LDS (0: read the global table)         */
  uint32_t kmax;           /* FTB: entry slots per pixel (8, 16 or 32)         */
  uint64_t estride;        /* FTB: pixels per entry slot plane                 */
  uint32_t composite;      /* VFR_COMPOSITE_* (both walks)                     */
  uint32_t psi_table;      /* FTB: psi from the table (VFR_RENDER_PSI_TABLE)     */
  uint32_t* vflags;        /* per-view VFR_VFLAG_* words, indexed like views, or
                              null                                             */
  uint32_t lds_records;    /* FTB, one ray per thread, psi in registers, records
                              below 4 GiB: tet records staged in LDS per wave
                              (VFR_RENDER_LDS_RECORDS)                         */
  const float4* memo;      /* REF: face-plane memo, 2 float4 per (view, tet,
                              face): (z0, z1, z2, s0), (s1, s2, 0, 0); or null */
  uint32_t memo_tets;      /* tets per view in the memo                        */
  uint32_t tf_block;       /* FTB, one ray per thread, more TF slots than LDS
                              holds: each block stages the TFs of its rays'
                              views (ray list in view order)                   */
};

/* ---- launchers (vfr_kernels.hip) ---- */

int vfr_k_clear_u32(uint32_t* p, size_t n, void* stream);
int vfr_k_copy(const void* src, void* dst, size_t bytes, void* stream);
/* Batch launches cover views [v0, v0+nv) with a grid of (the most vertices
 * or external faces any of those views' meshes has) x nv. */
int vfr_k_xform_views(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_verts,
                      const float4* wvert, float4* pv, void* stream);
/* The face-plane memo of views [v0, v0+nv): z and s planes of every tet face
 * (face f = slots f, f+1, f+2 of the tet's vertex ids) from the views'
 * P-space vertices, one record of 2 float4 per (view, tet, face). */
int vfr_k_face_memo(const DevView* views, uint32_t v0, uint32_t nv, uint32_t n_tets, const uint4* tv,
                    const float4* pverts, float4* memo, void* stream);
int vfr_k_raster_ref(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_faces, const uint4* fvid,
                     const float4* pv, uint32_t* owner, void* stream);
/* kmax: entry slots per pixel; 0 only counts (cnt) without storing keys. */
/* vflags (may be null): VFR_VFLAG_OVERFLOW is or-ed into the word of each
 * view where a pixel's list overflows. */
int vfr_k_raster_ftb(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_faces, const uint4* fvid,
                     const float4* pv, uint32_t* cnt, uint64_t* entries,
                     uint32_t kmax, uint64_t estride, unsigned long long* stats, uint32_t* vflags,
                     void* stream);
/* *out = max(*out, cnt[i]) over n pixels (device counter). */
int vfr_k_max_u32(const uint32_t* cnt, size_t n, uint32_t* out, void* stream);
/* view_base: added to the view index of every ray written (the batch's views
 * sit at [view_base, view_base + n_views_total) of the array the walk reads;
 * `views` points at the first of them). */
int vfr_k_compact(const DevView* views, uint32_t n_views_total, uint32_t tile0, uint32_t ntiles,
                  uint32_t mode, const uint32_t* owner, uint32_t* cnt, uint64_t* entries, uint32_t kmax,
                  uint64_t estride, float4* rgba, uint32_t* steps, uint32_t* entry_out, uint64_t* rays,
                  uint32_t* n_rays, void* stream, uint32_t view_base = 0,
                  uint32_t* block_off = nullptr, uint64_t* tile_mask = nullptr);
/* Blocks of a compaction launch over ntiles tiles, the size block_off needs.
 * With block_off (and tile_mask: ntiles uint64 masks + ntiles uint32 views),
 * rays are written in tile order after a count pass and a scan; without, each
 * block appends its rays with one atomic, in arrival order. */
uint32_t vfr_compact_blocks(uint32_t ntiles);
int vfr_k_clear_view_pixels(const DevView* views, uint32_t v0, uint32_t nv, uint32_t max_pix, uint32_t* a,
                            void* stream);
/* occupancy: waves per SIMD of the persistent kernel variant (4, 5, 6:
 * register-capped variants; anything else: as compiled). */
/* proj: VFR_TRAV_* (the projections of the views the rays belong to). */
enum { VFR_TRAV_ORTHO = 0, VFR_TRAV_PERSP = 1, VFR_TRAV_MIXED = 2 };
/* Walk kernel variants beside VFR_MODE_REF / VFR_MODE_FTB: the REF walk and
 * the front-to-back walk with linear compositing (TravArgs.composite !=
 * VFR_COMPOSITE_PPI). */
#define VFR_MODE_REF_LINEAR 2u
#define VFR_MODE_FTB_LINEAR 3u
/* n_records: tet records in the arena the rays walk (from VFR_BIG_RECORDS on,
 * the kernels address records with 64-bit arithmetic). */
/* Walk kernels: one ray per thread (k_trav), the persistent kernel with lane
 * refill (k_trav_persist; grid_waves waves, `occupancy` variant), the wave
 * work queue (k_trav_wq; a.queue must be zero).  The last two take
 * front-to-back walks only. */
enum { VFR_WALK_THREAD = 0, VFR_WALK_REFILL = 1, VFR_WALK_QUEUE = 2 };
int vfr_k_traverse(uint32_t mode, uint32_t proj, uint64_t n_records, uint32_t occupancy, const TravArgs& a,
                   uint32_t max_rays, int kernel, uint32_t grid_waves, void* stream);
/* Waves per SIMD the persistent variant for `occupancy` reaches. */
int vfr_k_traverse_waves_per_simd(uint32_t occupancy, uint32_t* out);

/* host helpers shared by plans and the server (vfr_api.hip) */
struct vfr_ctx;
struct vfr_mesh;
struct vfr_view;
int vfr_make_view(const vfr_view& in, const float* bb_lo, const float* bb_hi, DevView* o);
int vfr_ctx_device(const vfr_ctx* c);
void* vfr_ctx_stream(const vfr_ctx* c);
/* The mesh's (caller tet, face) pairs of its external faces, on the host. */
const uint32_t* vfr_mesh_host_extf(const vfr_mesh* m);
/* Size of vfr_render_noalloc's device scratch area (one per device), and its
 * set-up: the area's device address on `device` (the current device), taken
 * once per device by vfr_ctx_create. */
#define VFR_SCRATCH_BYTES (128ull << 20)
int vfr_noalloc_init(int device);
int vfr_ctx_cus(const vfr_ctx* c);
const float4* vfr_ctx_tf(const vfr_ctx* c);
uint32_t vfr_ctx_n_tf(const vfr_ctx* c);
const float* vfr_ctx_psi(const vfr_ctx* c);
const vfr_ctx* vfr_mesh_ctx(const vfr_mesh* m);
const DevMesh* vfr_mesh_dev(const vfr_mesh* m);
void vfr_mesh_bbox(const vfr_mesh* m, float lo[3], float hi[3]);

/* Adds `add` to the tet index of the linked neighbour words of n records
 * (a mesh copied into a record arena at record `add`). */
int vfr_k_rebase_links(TetRec* rec, uint32_t n, uint32_t add, void* stream);

/* Rotates the n vertices in place by mv as rotateVolume does (mv == NULL:
 * no rotation), then, if asked, renormalizes them as normalizeVertices does;
 * lo/hi receive the resulting x, y, z bounds.  Synchronizes the stream. */
int vfr_k_rotate_normalize(float4* v, uint32_t n, const float* mv, int normalize, float lo[3], float hi[3],
                           void* stream);

/* Preprocessing on the device: uploads the caller's vertices and tets
 * (and .con/.extf contents when given), checks every id, normalizes the
 * vertices (normalizeVertices) when asked, builds connectivity and external
 * faces by sort-and-match unless given, the limits (buildLmt) unless h_lmt is
 * given, the Morton orderings and the face records.  bb_lo/bb_hi receive the
 * x, y, z bounds of the (normalized) vertices. */
int vfr_dev_build_mesh(DevMesh& m, const float* h_xyzs, const uint32_t* h_tets,
                       const uint32_t* h_con, const uint32_t* h_extf, uint32_t n_ext_given,
                       bool normalize, bool morton, const float* h_lmt, float bb_lo[3], float bb_hi[3],
                       void* stream);
#endif /* __HIPCC__ */

#endif
