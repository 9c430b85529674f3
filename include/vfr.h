/*
 * vfr.h -- C ABI of the MI355X-native VF-Ray renderer (libvfr.so).
 *
 * VF-Ray renders unstructured tetrahedral meshes by ray casting: visible
 * external faces seed one ray per pixel, each ray walks tet-to-tet through the
 * face adjacency and composites a partially pre-integrated segment per tet.
 *
 * Conventions
 *   - Every function returns an int status: VFR_OK (0) or a negative VFR_E*
 *     code; vfr_last_error() returns the text of the last failure on the
 *     calling thread.
 *   - Host buffers are caller-owned; device buffers are library-owned.
 *   - Ids in and out of the library are the caller's ids (the order of the
 *     arrays given to vfr_mesh_create); any internal reordering is hidden.
 *   - One vfr_ctx may be used by one host thread at a time.
 *
 * Reference interface each group replaces (file:line in the VF-Ray sources):
 *   mesh / connectivity : readOff+normalizeVertices+buildIncid+buildCon+
 *                         buildExtF+buildLmt  (vfRayPreComp.h:97-279,
 *                         vfRayPreComp.cc:29-145, 219-270, 424-501, 710-737,
 *                         808-853)
 *   transfer function   : readTF/buildTF (vfRayPreComp.cc:551-622)
 *   psi table           : psiGammaTable512.h:10-20 (float[512][512])
 *   render              : findVisibleFaces + host compaction +
 *                         projectVisibleFaces + vfRay, driven by
 *                         executeKernel1..3 (vfRayKernelCPU.h:61-80,
 *                         main.cc:74-224, vfRay.cu:319-568)
 */
#ifndef VFR_H
#define VFR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define VFR_OK 0
#define VFR_EINVAL (-1)   /* bad argument (null pointer, size, id)          */
#define VFR_ENOMEM (-2)   /* host or device allocation failed               */
#define VFR_EHIP (-3)     /* HIP runtime error                              */
#define VFR_EIO (-4)      /* file missing, unreadable or malformed          */
#define VFR_ESTATE (-5)   /* object not in a state that allows the call     */
#define VFR_EBUDGET (-6)  /* request exceeds the configured HBM budget      */
#define VFR_ENODEV (-7)   /* no usable gfx950 device                        */
#define VFR_EMESH (-8)    /* mesh violates an assumption (see message)      */

const char* vfr_last_error(void);
/* Library version string, e.g. "vfr 0.1 gfx950". */
const char* vfr_version(void);

/* --------------------------------------------------------------- context */
typedef struct vfr_ctx vfr_ctx;
typedef struct vfr_mesh vfr_mesh;
typedef struct vfr_plan vfr_plan;
typedef struct vfr_server vfr_server;

/* Opens `device` (HIP ordinal), uploads the closed-form psi table. */
int vfr_ctx_create(int device, vfr_ctx** out);
void vfr_ctx_destroy(vfr_ctx* ctx);

/* Replaces the context's 512x512 psi table (row = gamma_front index,
 * column = gamma_back index, as psiGammaTable512.h:12-20).  NULL restores the
 * table generated from the closed form. */
int vfr_psi_load(vfr_ctx* ctx, const float* table512x512);
/* Copies the context's psi table to host (512*512 floats). */
int vfr_psi_read(vfr_ctx* ctx, float* out512x512);
/* The closed-form table on host, no device needed (512*512 floats). */
int vfr_psi_generate(float* out512x512);

/* Streams a plain device-to-device copy of `bytes` `iters` times and
 * returns the achieved read+write bandwidth in GB/s: the device's practical
 * HBM ceiling, and the yardstick for the profiler's byte counters.  No
 * counterpart in the reference. */
int vfr_bandwidth_probe(vfr_ctx* ctx, uint64_t bytes, int iters, double* gbps);

/* Registers a 256-entry RGBA transfer function (1024 floats, r g b a per
 * entry, as the .tf file, vfRayPreComp.cc:551-579).  Returns its id. */
int vfr_tf_create(vfr_ctx* ctx, const float rgba[1024], int* tf_id);

/* ------------------------------------------------------------------ mesh */
#define VFR_MESH_NORMALIZE 0x1u  /* apply normalizeVertices on upload      */
#define VFR_MESH_MORTON 0x2u     /* reorder tets along a Z-curve of their
                                     centroids, vertices in the order those
                                     tets first use them                   */

typedef struct vfr_mesh_info {
  uint32_t n_verts, n_tets, n_ext_faces;
  float max_edge_length, max_z, min_z; /* the .lmt triple                 */
  uint64_t device_bytes;               /* resident footprint              */
  double build_ms;                     /* device preprocessing time       */
} vfr_mesh_info;

/* Uploads a mesh: xyzs = nV*(x,y,z,s) floats, tets = nT*4 vertex ids.
 * con (nT*4, neighbour per face or the tet itself on external faces, the
 * .con body) and extf (n_ext*2, (tet, face) pairs, the .extf body) may be
 * given (read from cache files) or NULL, in which case they are built on the
 * device by sort-and-match over packed face keys.  lmt (3 floats) may be
 * given or NULL (built).  Flags: VFR_MESH_*. */
int vfr_mesh_create(vfr_ctx* ctx, const float* xyzs, uint32_t n_verts,
                    const uint32_t* tets, uint32_t n_tets,
                    const uint32_t* con, const uint32_t* extf, uint32_t n_ext,
                    const float* lmt, uint32_t flags, vfr_mesh** out);
void vfr_mesh_destroy(vfr_mesh* mesh);
int vfr_mesh_get_info(const vfr_mesh* mesh, vfr_mesh_info* out);
/* Copies connectivity back in caller ids: con_out nT*4 (may be NULL),
 * extf_out n_ext*2 (may be NULL), lmt_out 3 floats (may be NULL). */
int vfr_mesh_connectivity(const vfr_mesh* mesh, uint32_t* con_out,
                          uint32_t* extf_out, float* lmt_out);
/* Copies the (possibly normalized) vertices back, caller order, nV*4. */
int vfr_mesh_vertices(const vfr_mesh* mesh, float* xyzs_out);
/* The CUDA viewer's re-render step (vfRay.cu:618-640): rotateVolume
 * (vfRayPreComp.cc:886-906, mv read row-major) on the resident vertices,
 * then, with VFR_MESH_NORMALIZE, normalizeVertices (:84-145), both on the
 * device and bit-exact with the host functions below.  Connectivity and the
 * limits (.lmt) are kept, as the viewer keeps them (its maxEdgeLength goes
 * stale once the renormalization rescales the mesh; REF mode reproduces
 * that).  Plans created before the call return VFR_ESTATE until created
 * again; servers hold their own copy of the mesh. */
int vfr_mesh_rotate(vfr_mesh* mesh, const float mv[16], uint32_t flags);

/* ---------------------------------------------------------------- render */
#define VFR_PROJ_ORTHO 0u       /* X,Y,Z = A*(x,y,z)+b, rays march -Z      */
#define VFR_PROJ_PERSPECTIVE 1u /* pinhole camera                          */

/* Compositing modes.  REF reproduces vfRayKernelCPU.cc:448-482 (segment put
 * in front of the accumulation, no early termination) and the reference's
 * per-face pixel loop semantics (last visible face in order wins a pixel,
 * a column stops at an entry scalar outside [0,1]).  FTB composites front to
 * back from the eye with early termination and re-enters through concave
 * boundaries. */
#define VFR_MODE_REF 0u
#define VFR_MODE_FTB 1u

typedef struct vfr_view {
  uint32_t width, height;
  uint32_t projection;  /* VFR_PROJ_*                                      */
  int32_t tf_id;        /* transfer function id                            */
  float ortho[12];      /* row-major 3x4 affine for VFR_PROJ_ORTHO         */
  float eye[3];         /* VFR_PROJ_PERSPECTIVE                            */
  float target[3];
  float up[3];
  float fov_y_deg;      /* spans the image height; pixels are square, so the
                           width spans tan(fov_y/2) * width / height     */
} vfr_view;

/* Schedules. */
#define VFR_SCHED_PER_VIEW 0u  /* one K1/K2/K3 launch set per view, frame sync */
#define VFR_SCHED_BATCHED 1u   /* all views in one persistent work-queue launch */

typedef struct vfr_render_opts {
  uint32_t mode;        /* VFR_MODE_*                                      */
  uint32_t schedule;    /* VFR_SCHED_*                                     */
  float term_alpha;     /* FTB early termination (0.99 default if <= 0)    */
  uint32_t waves_per_cu;/* persistent kernel: resident waves per CU (0 = 16) */
  uint32_t refill;      /* persistent kernel: idle lanes that trigger a wave
                           refill (0 = 32)                                 */
  uint32_t use_graph;   /* capture the batched launch set in a hipGraph    */
  uint32_t max_entries; /* FTB entry-list slots per pixel: 8, 16 or 32; 0 =
                           the fewest that hold every pixel's entry faces
                           (counting pass at plan creation; a pixel with more
                           than 32 makes vfr_plan_create fail, VFR_EMESH)   */
  uint32_t flags;       /* VFR_RENDER_*                                    */
  uint32_t occupancy;   /* persistent kernel variant: waves per SIMD (4, 5,
                           6: register-capped; other values: the compiler's
                           choice); tuned by vfr_plan_autotune            */
  uint32_t composite;   /* VFR_COMPOSITE_* (both modes)                    */
} vfr_render_opts;

/* Segment compositing.  PPI: partial pre-integration through the psi table
 * (the reference's USEPARTIALPREINT 1 build, vfRayKernelCPU.cc:448-482).
 * LINEAR: the USEPARTIALPREINT 0 build's linear integration of 0..255
 * colours, truncated to integers and clamped at 255 every step, dz in object
 * units (vfRayKernel.cu:738-806 as the reference builds it, USEOPACITY
 * unset: no opacity, no early termination; the CPU version at
 * vfRayKernelCPU.cc:484-532 never stores its colour and renders black).
 * LINEAR_OPACITY adds the USEOPACITY terms: transparency 1 - opacity, the
 * dz^2 correction, opacity accumulation, termination past 1, and the image
 * byte color * opacity.  rgba holds the image byte / 255 and the opacity (1
 * for LINEAR).  In REF mode the linear modes follow the reference's walk; in
 * FTB mode the same segment formula runs front to back on the FTB walk
 * (perspective views, re-entry through concave boundaries) with dz the
 * segment's length in object units, the opacity stop as the only early
 * termination (term_alpha does not apply), on the one-ray-per-thread kernel
 * (the persistent and wave-queue flags are ignored), and psi is not used. */
#define VFR_COMPOSITE_PPI 0u
#define VFR_COMPOSITE_LINEAR 1u
#define VFR_COMPOSITE_LINEAR_OPACITY 2u

/* Batched FTB renders run one ray per thread over the device-built list of
 * active rays of all views (the hardware dispatcher refills the CUs wave by
 * wave) unless this flag asks for the persistent work-queue kernel (a fixed
 * grid of waves that claims rays from the list and refills idle lanes). */
#define VFR_RENDER_PERSISTENT 0x1u
/* Batched front-to-back renders (without a graph) run all preprocessing and
 * then one walk launch over every view, unless this flag asks for view
 * chunks whose preprocessing overlaps the walk of the chunks before them on
 * a second stream. */
#define VFR_RENDER_PIPELINED 0x2u
/* Front-to-back walks evaluate each segment's psi integral in registers
 * (one exponential times a 12-term series; truncation error below 4.1e-8 for
 * the optical depths a segment reaches) unless this flag asks for the
 * 512x512 table lookup (nearest entry of the
 * two gammas, the reference's vfRayKernelCPU.cc:454-470 convention; the
 * table is the one vfr_psi_load sets).  REF mode always uses the table. */
#define VFR_RENDER_PSI_TABLE 0x4u
/* Front-to-back walks use the 64-bit-address kernel variant that meshes of
 * 2^27 tets or more need, whatever the mesh size (for testing that variant
 * on small meshes; the results are the same). */
#define VFR_RENDER_WIDE 0x8u
/* Batched front-to-back walks run as a wave work queue: a grid that fills
 * the device, each wave taking the next 64 rays of the list with one atomic
 * until the list is drained (no refill inside a wave; a wave that is done
 * moves on without waiting for its block).  Not with VFR_RENDER_PERSISTENT. */
#define VFR_RENDER_WAVE_QUEUE 0x10u
/* The one-ray-per-thread front-to-back walk (psi in registers, meshes below
 * 2^27 tets) with tet records staged in LDS per wave: per step the lanes of
 * a wave that move into the same tet load its 32-byte record once (one owner
 * lane per tet, elected through a 64-slot table in LDS) and the others read
 * it from LDS.  Same results; other walks ignore the flag. */
#define VFR_RENDER_LDS_RECORDS 0x20u
/* REF-mode walks read each exit face's z and s planes from a table of the
 * planes of every tet face, per view, instead of solving them per step: the
 * GPU form of the reference CPU walk's face-plane memo
 * (vfRayKernelCPU.cc:234-245, 369-435), which keeps solved planes keyed by
 * (tet, face) so that later rays of the same face reuse them.  The table
 * (32 bytes per tet face per view) is filled after the vertex transform of
 * each render with the walk's own arithmetic, so the results are the same
 * bits.  Front-to-back walks ignore the flag (they solve no face planes), and
 * so does vfr_render_noalloc (no room for the table in its scratch area). */
#define VFR_RENDER_FACE_MEMO 0x40u

typedef struct vfr_frame_out {
  float* rgba;          /* W*H*4, may be NULL                              */
  uint8_t* rgb8;        /* W*H*3 PPM body (row = pix.y), may be NULL       */
  uint32_t* entry_ext;  /* W*H, ext-face index or 0xFFFFFFFF, may be NULL  */
  uint32_t* entry_tet;  /* W*H, caller tet id or 0xFFFFFFFF, may be NULL   */
  uint32_t* steps;      /* W*H, tets visited per ray, may be NULL          */
} vfr_frame_out;

typedef struct vfr_stats {
  uint64_t rays;          /* sum over views of W*H                          */
  uint64_t active_rays;   /* rays with an entry face                        */
  uint64_t steps;         /* tets visited (sum over rays)                   */
  uint64_t anomalies;     /* rays stopped with no exit face found           */
  double ms_total;        /* HIP-event time of the whole render             */
  double ms_visibility;   /* K1+K2 (visibility, projection, entry lists)    */
  double ms_traversal;    /* K3                                             */
  double ms_frame_p50;    /* per-view frame latency (PER_VIEW schedule)     */
  double ms_frame_p99;
} vfr_stats;

/* A plan owns every device buffer a render needs; executing it allocates
 * nothing.  Views are copied. */
int vfr_plan_create(vfr_ctx* ctx, const vfr_mesh* mesh, const vfr_view* views,
                    int n_views, const vfr_render_opts* opts, vfr_plan** out);
/* VFR_EMESH (whether or not stats is given) when entry faces were dropped
 * (a pixel with more entry faces than max_entries) or a ray met a neighbour
 * whose faces do not link back (non-manifold mesh). */
int vfr_plan_execute(vfr_plan* plan, vfr_stats* stats);
int vfr_plan_read(vfr_plan* plan, int view, vfr_frame_out* out);
void vfr_plan_destroy(vfr_plan* plan);
/* Per-view frame latency (ms, HIP events) of the last PER_VIEW execute. */
int vfr_plan_frame_times(const vfr_plan* plan, double* out_ms);

/* Front-to-back plans: times the one-ray-per-thread traversal (with and
 * without LDS-staged records), the wave work queue and candidate
 * persistent-kernel schedules (kernel occupancy variant x waves per CU x
 * refill threshold) on the plan itself and keeps the fastest walk, then
 * keeps the pipelined execution if it is faster end to end.  REF plans: times
 * the whole render with and without the face-plane memo and keeps the faster.
 * Returns the choice (flags holds VFR_RENDER_WAVE_QUEUE /
 * VFR_RENDER_PERSISTENT / VFR_RENDER_LDS_RECORDS / VFR_RENDER_PIPELINED /
 * VFR_RENDER_FACE_MEMO) and the time it chose by (the walk's; the whole
 * render's for REF plans).  Results are cached per (mode, tets, rays, views,
 * psi variant) in the process.  Batched plans only. */
int vfr_plan_autotune(vfr_plan* plan, uint32_t* flags, uint32_t* occupancy,
                      uint32_t* waves_per_cu, uint32_t* refill, double* best_ms);

/* create + execute + read + destroy. outs has n_views entries. */
int vfr_render(vfr_ctx* ctx, const vfr_mesh* mesh, const vfr_view* views,
               int n_views, const vfr_render_opts* opts, vfr_frame_out* outs,
               vfr_stats* stats);

/* vfr_render without any allocation: no hipMalloc/hipFree, no host heap.
 * No counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The per-pixel and per-view buffers live in a fixed device
 * scratch area of 128 MiB per device (VFR_SCRATCH_BYTES in the sources),
 * which the library holds from its first context on that device until the
 * process ends; its layout is the views, counters, per-pixel count, ray,
 * RGBA, step and entry buffers, the P-space vertices (16 B x vertices x
 * views) and the entry keys (8 B x slots x pixels).  A call that would need
 * more returns VFR_EBUDGET with the sizes in the message.  Every call on one
 * device shares that device's scratch area (each call clears what must start
 * at zero): calls that overlap on one device race.  Renders one batched
 * launch set on the context's stream and waits for it; opts as vfr_render,
 * except use_graph and VFR_RENDER_PIPELINED (VFR_EINVAL); rgb8 outputs need
 * rgba outputs beside them; stats carry the host wall time in ms_total. */
int vfr_render_noalloc(vfr_ctx* ctx, const vfr_mesh* mesh, const vfr_view* views,
                       int n_views, const vfr_render_opts* opts, vfr_frame_out* outs,
                       vfr_stats* stats);

/* --------------------------------------------------------------- serving */
/* A render server keeps meshes resident on the device and renders a stream of
 * requests with continuous batching.  Two schedules:
 *
 *  VFR_SERVE_SLOTS     requests queue on arrival; whenever one of `streams`
 *                      slots is free, the queued requests that fit the slot's
 *                      share of the HBM budget (front of the queue first, up
 *                      to max_batch) are rendered together in one batched
 *                      launch set on that slot's stream.
 *  VFR_SERVE_PIPELINE  requests are admitted as soon as their buffers fit in
 *                      the server's shared arenas: their preprocessing
 *                      (vertex transform, visibility, entry lists, ray
 *                      compaction) runs at once on a preprocessing stream and
 *                      appends their rays to the open walk list, while up to
 *                      `streams` earlier walk lists walk; whenever a walk
 *                      stream is free the open list is walked in one launch,
 *                      however many requests joined it meanwhile.
 *
 * No counterpart in the reference (one 512x512 frame per process,
 * main.cc:242-380). */
#define VFR_SERVE_SLOTS 0u
#define VFR_SERVE_PIPELINE 1u
typedef struct vfr_server_cfg {
  uint64_t hbm_budget;  /* bytes for per-request buffers, all slots (0 = 16 GiB) */
  uint32_t streams;     /* slots = concurrent batches (0 = 2)                    */
  uint32_t max_batch;   /* requests per batch (0 = 256)                          */
  float term_alpha;     /* early termination (0 = 0.99)                          */
  uint32_t flags;       /* VFR_RENDER_PERSISTENT: persistent traversal kernel;
                           VFR_RENDER_PSI_TABLE: psi from the table              */
  uint32_t occupancy;   /* its schedule (0 = defaults, as vfr_render_opts)       */
  uint32_t waves_per_cu;
  uint32_t refill;
  uint32_t composite;   /* VFR_COMPOSITE_*, for every request (0 = PPI); the
                           linear modes use the one-ray-per-thread walk      */
  uint64_t max_batch_pixels; /* pixels per batch (slots) or per walk list
                           (pipeline); 0 = as the buffers allow; a request
                           larger than the cap still runs, alone             */
  uint32_t schedule;    /* VFR_SERVE_SLOTS (0) or VFR_SERVE_PIPELINE; with the
                           pipeline `streams` is the number of walks in flight */
} vfr_server_cfg;

typedef struct vfr_request {
  int32_t mesh;         /* index into the server's meshes                        */
  uint32_t keep_image;  /* copy the RGBA result back for vfr_server_wait         */
  vfr_view view;        /* perspective or orthographic, with its tf_id           */
} vfr_request;

typedef struct vfr_result {
  int32_t status;       /* 0 queued, 1 running, 2 done, < 0 error code: VFR_EMESH
                           when the view has a pixel with more than 32 entry
                           faces or a ray met a non-manifold face (no image) */
  uint32_t width, height;
  uint32_t batch_size;  /* requests rendered in the same batch                   */
  double latency_ms;    /* arrival -> result complete on the device              */
  double queue_ms;      /* arrival -> admitted into a batch                      */
} vfr_result;

typedef struct vfr_server_report {
  uint64_t requests, rays, active_rays, steps, anomalies, batches;
  double makespan_s;    /* first arrival -> last completion                      */
  double mrays_per_s;   /* rays / makespan                                       */
  double latency_p50_ms, latency_p90_ms, latency_p99_ms, latency_max_ms;
  double mean_batch;
} vfr_server_report;

/* meshes stay owned by the caller and must outlive the server; their
 * device data is copied into the server's resident arenas. */
int vfr_server_create(vfr_ctx* ctx, const vfr_server_cfg* cfg, const vfr_mesh* const* meshes,
                      int n_meshes, vfr_server** out);
/* Queues a request (arrival = now); VFR_EBUDGET if it can never fit a slot. */
int vfr_server_submit(vfr_server* s, const vfr_request* req, uint64_t* ticket);
int vfr_server_poll(vfr_server* s, uint64_t ticket, vfr_result* out);
/* Blocks until done; copies the image (W*H*4 floats) if requested and kept,
 * then forgets the ticket. */
int vfr_server_wait(vfr_server* s, uint64_t ticket, vfr_result* out, float* rgba_out);
/* Load generator: submits reqs[i] at arrival offsets_s[i] seconds after the
 * call, waits for all, and reports throughput and latency percentiles; the
 * first request error, if any, is returned after every request is reaped. */
int vfr_server_replay(vfr_server* s, const vfr_request* reqs, const double* offsets_s, int n,
                      vfr_server_report* report);
void vfr_server_destroy(vfr_server* s);

/* --------------------------------------------------------------- formats */
/* Text formats of vfRayPreComp.cc, byte-identical output.  Readers allocate
 * with malloc; free with vfr_free. */
void vfr_free(void* p);
int vfr_read_off(const char* path, float** xyzs, uint32_t* n_verts,
                 uint32_t** tets, uint32_t* n_tets);           /* :29-80   */
int vfr_write_off(const char* path, const float* xyzs, uint32_t n_verts,
                  const uint32_t* tets, uint32_t n_tets);
void vfr_normalize_vertices(float* xyzs, uint32_t n_verts);    /* :84-145  */
void vfr_rotate_volume(float* xyzs, uint32_t n_verts, const float mv[16]); /* :886-906 */
/* The viewer's modelview for its rotation angles (vfRay.cu:618-622):
 * Rx(rotx) Ry(roty), column-major, computed in double and rounded. */
void vfr_gl_rotation(float rotx_deg, float roty_deg, float mv[16]);
int vfr_read_con(const char* path, uint32_t n_tets, uint32_t** con,
                 uint32_t* n_ext);                             /* :375-419 */
/* readIncid (:153-212): the per-vertex incident-tet lists of a .incid file as
 * CSR arrays (offsets: n_verts + 1 entries, tet_ids), malloc'ed. */
int vfr_read_incid(const char* path, uint32_t n_verts, uint32_t** offsets,
                   uint32_t** tet_ids);
/* buildCon (:424-501) over such lists, on the host: con_out (n_tets*4) and
 * the number of external faces, exactly as the reference derives them from
 * the lists it read (the GPU sort-and-match of vfr_mesh_create gives the
 * same result for the lists buildIncid writes). */
int vfr_con_from_incid(const uint32_t* tets, uint32_t n_tets, uint32_t n_verts,
                       const uint32_t* offsets, const uint32_t* tet_ids,
                       uint32_t* con_out, uint32_t* n_ext_out);
int vfr_write_con(const char* path, const uint32_t* con, uint32_t n_tets,
                  uint32_t n_ext);                             /* :507-543 */
int vfr_read_extf(const char* path, uint32_t n_ext, uint32_t** extf); /* :668-705 */
int vfr_write_extf(const char* path, const uint32_t* extf, uint32_t n_ext); /* :743-774 */
int vfr_read_tf(const char* path, float rgba_out[1024]);       /* :551-590 */
int vfr_write_tf(const char* path, const float rgba[1024]);    /* :628-660 */
void vfr_default_tf(float rgba_out[1024]);                     /* :595-622 */
/* cleanTF (cleanTF.cc:14-78): legacy TF file (7 header lines, 256 lines
 * "index r g b a") to the .tf text. */
int vfr_clean_tf(const char* legacy_path, const char* tf_path);
int vfr_read_lmt(const char* path, float lmt_out[3]);          /* :782-803 */
int vfr_write_lmt(const char* path, const float lmt[3]);       /* :859-880 */
void vfr_build_lmt(const float* xyzs, uint32_t n_verts, const uint32_t* tets,
                   uint32_t n_tets, float lmt_out[3]);         /* :808-853 */
/* buildIncid + writeIncid (vfRayPreComp.cc:219-332): the .incid cache the
 * reference writes on its way to the .con file. */
int vfr_write_incid(const char* path, const uint32_t* tets, uint32_t n_tets,
                    uint32_t n_verts);
int vfr_write_ppm(const char* path, const uint8_t* rgb, uint32_t w,
                  uint32_t h);                                 /* :915-934 */

/* ------------------------------------------------------------- synthetic */
/* Seeded meshes of the benchmark recipes (hex grid of n^3 cells on [0,1]^3,
 * 6 Freudenthal tets per cell, jitter, Gaussian scalar field).
 * variant 0: plain grid; variant 1: concave (graded radial map + voids).
 * The arrays are malloc'ed; free them with vfr_free. */
int vfr_synth_grid(uint32_t n, uint64_t seed, uint32_t variant, float** xyzs,
                   uint32_t* n_verts, uint32_t** tets, uint32_t* n_tets);
/* 256-entry RGBA TF, piecewise linear through 6 seeded control points,
 * quantized through the 6-significant-digit .tf text. */
void vfr_synth_tf(uint64_t seed, float rgba_out[1024]);
/* Rotation that maps unit direction d to -z, applied about (0.5,0.5,0.5):
 * the reference renders along -z, so this bakes a view into the .off. */
void vfr_synth_rotate(float* xyzs, uint32_t n_verts, const float d[3]);

#ifdef __cplusplus
}
#endif
#endif /* VFR_H */
