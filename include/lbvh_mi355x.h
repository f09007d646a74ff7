/*
 * lbvh_mi355x.h — C ABI of the MI355X (gfx950) LBVH engine.
 *
 * Plain C: opaque handles, POD option structs, integer status codes, no
 * exceptions across the boundary.  Each entry point names the tay10r/lbvh
 * interface it replaces (file:line in the reference repository).
 *
 * Data layouts at the boundary are the reference's own:
 *   triangle  48 B  lbvh::triangle<float> {p0, e0 = a - b, e1 = c - a, normal}
 *                   (lbvh.h:1959-1983)
 *   aabb      24 B  lbvh::aabb<float> {min xyz, max xyz} (lbvh.h:205-212)
 *   node      32 B  lbvh::node<float> {aabb box; u32 left; u32 right}, bit 31
 *                   of left/right marks a leaf, root = node 0 (lbvh.h:220-255)
 *   ray       32 B  lbvh::ray<float> {pos, dir, tmin, tmax} (lbvh.h:387-401)
 *
 * Threading: a context is used by one host thread at a time; separate
 * contexts may be used concurrently.  Work is queued on the context's stream
 * (its own, or one installed with lbvh_ctx_set_stream).  Entry points taking
 * host pointers copy in and out and return after the work has finished;
 * the *_device variants take device pointers, borrow them, and return after
 * queuing (call lbvh_ctx_synchronize before reading results).
 */
#ifndef LBVH_MI355X_H
#define LBVH_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lbvh_status
{
  LBVH_OK = 0,
  LBVH_E_INVALID = 1,  /* bad argument: n < 2, null pointer, bad option */
  LBVH_E_OOM = 2,      /* device allocation failed or memory cap exceeded */
  LBVH_E_HIP = 3,      /* a HIP runtime call failed */
  LBVH_E_BUSY = 4,     /* no free slot / request rejected for now */
  LBVH_E_NODEVICE = 5, /* no usable gfx950 device */
  LBVH_E_TIMEOUT = 6   /* a kernel's bounded inter-workgroup wait expired */
} lbvh_status;

typedef struct lbvh_ctx lbvh_ctx;
typedef struct lbvh_bvh lbvh_bvh;

/* Leaf numbering.  REFERENCE reproduces lbvh.h byte for byte: a leaf index
 * is the primitive's position in Morton order, and boxes are fitted from
 * primitives[position] (lbvh.h:1728-1729, 1855, 1862).  PRIMITIVE_ID stores
 * the id of the primitive at that position (same topology). */
enum
{
  LBVH_LEAF_REFERENCE = 0,
  LBVH_LEAF_PRIMITIVE_ID = 1
};

/* Closest-hit intersector.  REFERENCE is lbvh::triangle_intersector
 * (lbvh.h:2055-2075) with the traverser's first-hit-wins tie rule
 * (lbvh.h:1909, 1920); REFERENCE_LOWID breaks equal-distance ties toward
 * the lowest primitive id; WATERTIGHT uses the watertight test of Woop,
 * Benthin and Wald (JCGT 2013) with lowest-id ties. */
enum
{
  LBVH_ISECT_REFERENCE = 0,
  LBVH_ISECT_REFERENCE_LOWID = 1,
  LBVH_ISECT_WATERTIGHT = 2
};

typedef struct lbvh_build_opts
{
  uint32_t code_bits;       /* 30 (u32 codes, as builder<float>) or 63 (u64) */
  uint32_t leaf_order;      /* LBVH_LEAF_* */
  uint32_t sort_digit_bits; /* 0 = tuned default; else 6..10 */
  uint32_t sort_tile;       /* 0 = tuned default; else keys per workgroup */
  uint32_t keep_codes;      /* keep sorted codes / ids for download */
  uint32_t sort_strategy;   /* 0 = tuned; 1 = onesweep (decoupled look-back);
                               2 = reduce-then-scan; 3 = hybrid (top digits
                               by reduce-then-scan, then each run of equal
                               top bits sorted by the full key in LDS) */
  uint32_t reserved[2];
} lbvh_build_opts;

typedef struct lbvh_bvh_info
{
  uint64_t n_prims;
  uint64_t n_nodes; /* n_prims - 1 */
  uint32_t code_bits;
  uint32_t leaf_order;
  const void* d_nodes;   /* device lbvh::node<float>[n_nodes] */
  const void* d_prims;   /* device triangles (NULL for AABB builds) */
  uint64_t device_bytes; /* device memory held by the handle */
  uint32_t sort_digit_bits; /* the sort configuration the build used */
  uint32_t sort_tile;
  uint32_t sort_strategy;   /* 1 onesweep, 2 reduce-then-scan, 3 hybrid */
  uint32_t sort_global_digits; /* hybrid: top digits sorted by global passes */
} lbvh_bvh_info;

/* One configuration timed by the context's autotuner: the sort of a build
 * (kind 0) or a closest-hit traversal (kind 1). */
typedef struct lbvh_tune_record
{
  uint32_t batch;    /* 1: a scene batch, 0: a single-scene build */
  uint32_t key_bits; /* sorted key width (builds) */
  uint32_t log2_n;   /* size bucket: ceil(log2(primitives)), or of rays */
  uint32_t sort_digit_bits; /* builds: the sort configuration */
  uint32_t sort_tile;
  uint32_t sort_strategy;
  uint32_t picked;   /* the configuration kept for this bucket */
  float ms;          /* device time of the whole build (traversal) with it */
  uint32_t scalar_bits; /* 32: builder<float> builds, 64: builder<double>
                           builds (tuned in buckets of their own) */
  uint32_t kind;     /* 0: a build's sort, 1: a traversal */
  uint32_t trace_reorder;        /* traversals: rays put in key order first */
  uint32_t trace_waves_per_simd; /* traversals: waves sharing a SIMD (8: as
                                    many as the registers allow) */
  uint32_t trace_mode;           /* traversals: the intersector mode tuned
                                    (LBVH_ISECT_*); each mode is tuned apart */
} lbvh_tune_record;

/* -- status / options ---------------------------------------------------- */

const char* lbvh_status_string(lbvh_status s);
void lbvh_build_opts_init(lbvh_build_opts* o); /* 30-bit, REFERENCE leaves */

/* -- context ---------------------------------------------------------------
 * Replaces the scheduler plugin of lbvh.h (single_thread_scheduler /
 * naive_thread_scheduler, lbvh.h:87-165): work runs on a HIP stream. */
lbvh_status lbvh_ctx_create(int device, lbvh_ctx** out);
void lbvh_ctx_destroy(lbvh_ctx* ctx);
lbvh_status lbvh_ctx_set_stream(lbvh_ctx* ctx, void* hip_stream);
lbvh_status lbvh_ctx_synchronize(lbvh_ctx* ctx);
/* Device timing of the following calls with HIP events on the context
 * stream: mode 0 off, 1 every phase, 2 one interval per call (no events
 * between the kernels of a call).  Setting a mode starts a new record. */
lbvh_status lbvh_ctx_set_timing(lbvh_ctx* ctx, int mode);
/* Milliseconds of the phases of the last timed call: for a build
 * [bounds+encode, sort, tree, fit], for a trace [trace], for a refit [fit]. */
lbvh_status lbvh_ctx_phase_ms(lbvh_ctx* ctx, float* ms, uint32_t cap,
                              uint32_t* count);
/* Phase milliseconds summed over every call of one kind (1 build, 2 trace,
 * 3 refit) since timing was last enabled; *calls receives the call count.
 * Events are recorded on the context stream without synchronizing, so a
 * timed loop runs back to back. */
lbvh_status lbvh_ctx_phase_stats(lbvh_ctx* ctx, int kind, float* sum_ms,
                                 uint32_t cap, uint32_t* count,
                                 uint32_t* calls);
const char* lbvh_ctx_last_error(const lbvh_ctx* ctx);
/* Autotuner (on by default).  A build whose options leave sort_digit_bits,
 * sort_tile and sort_strategy at 0 uses the sort configuration tuned for its
 * key width and size bucket; the first such build of a bucket times every
 * candidate configuration on its own data (the faster of two device-timed
 * runs each) and keeps the fastest.  lbvh_trace_closest[_device] does the
 * same per intersector mode and ray-count bucket (from 16,384 rays; one
 * untimed run, then the fastest of three per candidate) over the
 * traversal's configurations: rays traced as given or first put in the order
 * of a key (direction octant and a coarse position along the ray, so that a
 * wave's rays visit the same nodes), and the waves sharing a SIMD (8, 6 or
 * 4).  Results are identical for every configuration; only the speed
 * differs.  Disabled: fixed defaults (rays as given, 8 waves). */
lbvh_status lbvh_ctx_set_autotune(lbvh_ctx* ctx, int enable);
/* Copies up to cap records of every configuration timed so far; returns the
 * total number of records. */
uint32_t lbvh_ctx_tune_log(const lbvh_ctx* ctx, lbvh_tune_record* out,
                           uint32_t cap);
/* Device bytes held by the context's scratch workspace. */
uint64_t lbvh_ctx_workspace_bytes(const lbvh_ctx* ctx);

/* -- build -----------------------------------------------------------------
 * Replaces lbvh::builder<float>::operator()(primitives, count, converter)
 * (lbvh.h:342-345, 1793-1822) for P = triangle<float> with
 * C = triangle_aabb_converter<float> (lbvh.h:1988-2013). */
lbvh_status lbvh_build_triangles(lbvh_ctx* ctx, const float* tri48,
                                 uint64_t n, const lbvh_build_opts* opts,
                                 lbvh_bvh** out);
lbvh_status lbvh_build_triangles_device(lbvh_ctx* ctx, const float* d_tri48,
                                        uint64_t n,
                                        const lbvh_build_opts* opts,
                                        lbvh_bvh** out);
/* Same builder for any primitive whose converter was evaluated by the
 * caller: boxes24[i] = converter(primitives[i]) (lbvh.h:1289, 1453, 1855). */
lbvh_status lbvh_build_aabbs(lbvh_ctx* ctx, const float* boxes24, uint64_t n,
                             const lbvh_build_opts* opts, lbvh_bvh** out);
lbvh_status lbvh_build_aabbs_device(lbvh_ctx* ctx, const float* d_boxes24,
                                    uint64_t n, const lbvh_build_opts* opts,
                                    lbvh_bvh** out);

/* Refit: re-runs the box fit (lbvh.h:1827-1866) over new primitive data on
 * the unchanged topology.  The reference has no refit entry point (the fit
 * is protected, lbvh.h:347-352); the result equals a fresh fit. */
lbvh_status lbvh_refit_triangles(lbvh_ctx* ctx, lbvh_bvh* bvh,
                                 const float* tri48);
lbvh_status lbvh_refit_triangles_device(lbvh_ctx* ctx, lbvh_bvh* bvh,
                                        const float* d_tri48);

/* -- allocation-free build --------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: lbvh_build_triangles_device's build with no allocation
 * of any kind.  Every scratch buffer (centroids, keys, ids, sort counts, tree
 * segments and tickets, and the parent links the nodes are built beside) is
 * carved from one fixed device array that each device holds for the life of
 * the process (24 MiB, set up when a context is first created on the
 * device).  Inputs: d_tri48, n device triangles (16-byte aligned), n in
 * [2, LBVH_NOALLOC_MAX_PRIMS].  Output: the (n - 1) lbvh::node<float> records
 * of the tree in d_nodes_out (device memory, 32-byte aligned, (n - 1) x 32
 * bytes), byte-identical to lbvh_build_triangles_device's node array for the
 * same options.  keep_codes must be 0; the onesweep sort is rejected; with
 * the sort fields at 0 the context's tuned choice for the size is used when
 * one exists (no tuning runs), else a fixed default.  Larger n is rejected
 * with LBVH_E_INVALID.  Returns after queuing on the context stream.  Every
 * call on a device shares that device's scratch, so calls that overlap on one
 * device (contexts with different streams) race; calls on one stream do not. */
#define LBVH_NOALLOC_MAX_PRIMS 262144u
lbvh_status lbvh_build_triangles_noalloc(lbvh_ctx* ctx, const float* d_tri48,
                                         uint64_t n,
                                         const lbvh_build_opts* opts,
                                         void* d_nodes_out);

/* -- the built tree --------------------------------------------------------
 * Replaces lbvh::bvh<float> (lbvh.h:259-299): nodes stay resident on the
 * device; download yields the reference's node array. */
lbvh_status lbvh_bvh_get_info(const lbvh_bvh* bvh, lbvh_bvh_info* out);
lbvh_status lbvh_bvh_download_nodes(lbvh_ctx* ctx, const lbvh_bvh* bvh,
                                    void* dst, uint64_t cap_bytes);
/* Sorted Morton codes (u32 for 30-bit, u64 for 63-bit), the primitive ids in
 * sorted order (space_filling_curve entries after sort, lbvh.h:1104-1134) and
 * the centroid bounds (6 floats).  Needs opts.keep_codes; any pointer may be
 * NULL. */
lbvh_status lbvh_bvh_download_codes(lbvh_ctx* ctx, const lbvh_bvh* bvh,
                                    void* codes, uint32_t* ids,
                                    float* bounds6);
void lbvh_bvh_destroy(lbvh_bvh* bvh);

/* -- builder<double> --------------------------------------------------------
 * Replaces lbvh::builder<double>::operator()(primitives, count, converter)
 * (lbvh.h:342-345, 549-558, 1793-1822 with scalar_type = double) for any
 * primitive whose converter the caller evaluated: boxes48[i] =
 * converter(primitives[i]) as lbvh::aabb<double> {min xyz, max xyz}
 * (lbvh.h:205-212).  Codes are u64 (morton_domain<8>, lbvh.h:1322-1326) and
 * all box arithmetic is FP64.  Output: n - 1 lbvh::node<double> records of
 * 64 bytes {aabb<double> box; u64 left; u64 right}, bit 63 of left/right
 * marks a leaf, root = node 0 (lbvh.h:220-255).  opts->leaf_order and the
 * sort fields apply; code_bits and keep_codes are ignored (the code width
 * follows the scalar type, as in the reference); the sort is tuned per size
 * as for 63-bit FP32 builds.  n in [2, 2^31 - 1].  The device variant borrows
 * d_boxes48 and writes d_nodes_out, both device memory aligned to 16 bytes,
 * and returns after queuing; the host variant copies in and out and
 * returns when the nodes are in nodes_out. */
lbvh_status lbvh_build_aabbs_f64(lbvh_ctx* ctx, const double* boxes48,
                                 uint64_t n, const lbvh_build_opts* opts,
                                 void* nodes_out);
lbvh_status lbvh_build_aabbs_f64_device(lbvh_ctx* ctx,
                                        const double* d_boxes48, uint64_t n,
                                        const lbvh_build_opts* opts,
                                        void* d_nodes_out);

/* -- traversal -------------------------------------------------------------
 * Replaces lbvh::traverser<float, triangle<float>,
 * triangle_intersection<float>>::operator()(ray, triangle_intersector)
 * (lbvh.h:531-533, 1869-1942) applied to a batch of rays.  Outputs per ray:
 * prim (UINT32_MAX = miss, lbvh.h:366), t (+inf on a miss), uv (2 floats,
 * 0 on a miss).  uv may be NULL.  Every traversal entry point (single tree,
 * batch, FP64, server request) takes at most 2^31 - 1 rays per call. */
lbvh_status lbvh_trace_closest(lbvh_ctx* ctx, const lbvh_bvh* bvh,
                               const float* rays32, uint64_t nrays, int mode,
                               uint32_t* prim, float* t, float* uv);
lbvh_status lbvh_trace_closest_device(lbvh_ctx* ctx, const lbvh_bvh* bvh,
                                      const float* d_rays32, uint64_t nrays,
                                      int mode, uint32_t* d_prim, float* d_t,
                                      float* d_uv);

/* -- builder<double> traversal ----------------------------------------------
 * Replaces lbvh::traverser<double, triangle<double>,
 * triangle_intersection<double>>::operator()(ray, triangle_intersector)
 * (lbvh.h:531-533, 1869-1942, 2055-2075) applied to a batch of rays, over the
 * node<double> array lbvh_build_aabbs_f64 wrote for n triangles (n - 1
 * records of 64 bytes).  tri96: n triangle<double> records {p0, e0, e1,
 * normal} (12 doubles, lbvh.h:1960-1987); a leaf value v tests record v.
 * rays64: ray<double> records {pos xyz, dir xyz, tmin, tmax} (8 doubles,
 * lbvh.h:388-401).  All arithmetic is FP64 in the reference's order.
 * Outputs per ray: prim u64 (UINT64_MAX = miss, lbvh.h:360-366), t (+inf on
 * a miss), uv (2 doubles, 0 on a miss; may be NULL).  mode 0 = reference
 * (first hit found wins ties), 1 = lowest leaf value on equal distance.  n in
 * [2, 2^31 - 1].  nodes64 must be the node array of a tree over the n
 * records; a ray stops loading nodes after n visits (a tree never needs
 * more), so a corrupt array ends instead of looping.  The device variant
 * reads device memory (nodes, triangles, rays and uv 16-byte aligned, prim
 * and t 8-byte) and returns after queuing; the host variant copies in and out
 * and returns when the outputs are written. */
lbvh_status lbvh_trace_closest_f64(lbvh_ctx* ctx, const void* nodes64,
                                   uint64_t n, const double* tri96,
                                   const double* rays64, uint64_t nrays,
                                   int mode, uint64_t* prim, double* t,
                                   double* uv);
lbvh_status lbvh_trace_closest_f64_device(lbvh_ctx* ctx, const void* d_nodes64,
                                          uint64_t n, const double* d_tri96,
                                          const double* d_rays64,
                                          uint64_t nrays, int mode,
                                          uint64_t* d_prim, double* d_t,
                                          double* d_uv);

/* -- scene batches -----------------------------------------------------------
 * Many independent triangle scenes built in one segmented pass (per-scene
 * bounds and codes, one radix sort keyed by (scene, code), then tree and fit
 * over every scene at once).  Each scene's tree is the tree
 * lbvh::builder<float>::operator() (lbvh.h:1793-1822) builds for that scene
 * alone with 30-bit codes, as lbvh_build_triangles gives it.  The reference
 * has no batched builder: a batch replaces one builder call per scene.
 *
 * Scene s is triangles [prim_off[s], prim_off[s + 1]) of one array, with
 * prim_off[0] = 0, at least 2 triangles per scene and fewer than 2^31 in all.
 * Scene s's nodes start at node prim_off[s] - s of the batch's node array;
 * node indices, leaf values and hit ids are scene-local.  opts.code_bits
 * must be 30; sort_digit_bits / sort_tile / sort_strategy pick the sort
 * (0 = tuned); the onesweep strategy is rejected (its look-back state
 * belongs to one build, so a replayed graph could not reuse it).
 *
 * A batch owns every device buffer its rebuilds, refits and traversals need,
 * so lbvh_batch_rebuild and lbvh_batch_refit allocate nothing and do not
 * synchronize.  With use_graph != 0 the work is captured into a HIP graph on
 * first use (per stream and primitive pointer) and replayed afterwards. */
typedef struct lbvh_batch lbvh_batch;

typedef struct lbvh_batch_info
{
  uint32_t n_scenes;
  uint32_t leaf_order;
  uint64_t n_prims;
  uint64_t n_nodes;         /* n_prims - n_scenes */
  uint32_t sort_digit_bits; /* the sort configuration in use */
  uint32_t sort_tile;
  uint32_t sort_passes;
  uint32_t sort_strategy;   /* 2 reduce-then-scan, 3 hybrid */
  uint64_t device_bytes;  /* device memory held by the handle */
  const void* d_nodes;    /* device lbvh::node<float>[n_nodes] */
  const void* d_prims;    /* device triangles */
  uint32_t sort_global_passes; /* radix passes over the whole array per rebuild
                                  (hybrid: 1 when the keys' global digits hold
                                  one digit of code bits and each scene is
                                  sorted by it in one scene-segmented pass) */
  uint32_t reserved;
} lbvh_batch_info;

/* prim_off: host array of n_scenes + 1 offsets. */
lbvh_status lbvh_build_batch(lbvh_ctx* ctx, const float* tri48,
                             const uint64_t* prim_off, uint32_t n_scenes,
                             const lbvh_build_opts* opts, lbvh_batch** out);
lbvh_status lbvh_build_batch_device(lbvh_ctx* ctx, const float* d_tri48,
                                    const uint64_t* prim_off,
                                    uint32_t n_scenes,
                                    const lbvh_build_opts* opts,
                                    lbvh_batch** out);
/* Rebuild every scene (same scene sizes) from the current triangles, or from
 * d_tri48 when not NULL (device memory; copied into owned storage, or
 * borrowed in place of the previous pointer). */
lbvh_status lbvh_batch_rebuild(lbvh_ctx* ctx, lbvh_batch* batch,
                               const float* d_tri48, int use_graph);
/* Refit every scene on its unchanged topology (the box fit of
 * lbvh.h:1827-1866; the reference has no refit entry point). */
lbvh_status lbvh_batch_refit(lbvh_ctx* ctx, lbvh_batch* batch,
                             const float* d_tri48, int use_graph);
lbvh_status lbvh_batch_get_info(const lbvh_batch* batch, lbvh_batch_info* out);
/* One scene's node array ((n_s - 1) x 32 B, as lbvh::bvh<float>), sorted
 * 30-bit codes, scene-local primitive ids in sorted order (needs
 * PRIMITIVE_ID leaves or keep_codes) and centroid bounds (6 floats).  Any
 * output may be NULL. */
lbvh_status lbvh_batch_download_scene(lbvh_ctx* ctx, const lbvh_batch* batch,
                                      uint32_t scene, void* nodes,
                                      uint64_t cap_bytes, uint32_t* codes,
                                      uint32_t* ids, float* bounds6);
void lbvh_batch_destroy(lbvh_batch* batch);
/* Closest hits for rays of every scene in one launch: rays
 * [ray_off[s], ray_off[s + 1]) go to scene s (n_scenes + 1 offsets, host
 * memory for lbvh_batch_trace_closest, device memory for the _device
 * variant); outputs as lbvh_trace_closest with scene-local ids. */
lbvh_status lbvh_batch_trace_closest(lbvh_ctx* ctx, const lbvh_batch* batch,
                                     const float* rays32,
                                     const uint64_t* ray_off, int mode,
                                     uint32_t* prim, float* t, float* uv);
lbvh_status lbvh_batch_trace_closest_device(lbvh_ctx* ctx,
                                            const lbvh_batch* batch,
                                            const float* d_rays32,
                                            const uint64_t* d_ray_off,
                                            uint64_t nrays, int mode,
                                            uint32_t* d_prim, float* d_t,
                                            float* d_uv);

/* -- serving ------------------------------------------------------------------
 * A request server for one GPU with continuous batching.  A request names a
 * scene -- a new one (its triangles), or a resident one, optionally with new
 * triangle data (a refit on the unchanged topology) -- and carries a ray
 * batch; its result is the closest hit of every ray, as lbvh_trace_closest
 * gives it on that scene.  Built trees stay resident.  Requests are admitted
 * in arrival order as in-flight slots and device memory (under
 * memory_cap_bytes) free up; idle resident scenes are evicted least recently
 * used first when a new scene needs room.  Admitted requests build, refit and
 * upload on several streams at once, and every request whose scene is ready
 * joins the next traversal launch, which covers all of their trees and rays.
 * Requests on one scene keep arrival order (a refit affects exactly the
 * requests submitted after it), so each result equals the result of serving
 * the requests one at a time.  The reference has no server; its traversal
 * (lbvh.h:1869-1942) and builder (lbvh.h:1793-1822) are what each request
 * runs.  Host buffers named by a request must stay valid until the request
 * is collected by lbvh_server_wait or lbvh_server_poll. */
typedef struct lbvh_server lbvh_server;

typedef struct lbvh_server_opts
{
  uint64_t memory_cap_bytes; /* device bytes for scenes and requests (288 GiB,
                                lowered to the device's free memory) */
  uint32_t lanes;            /* build / refit / upload streams (4) */
  uint32_t max_inflight;     /* admitted requests at once (64) */
  uint32_t code_bits;        /* 30 or 63 (63) */
  uint32_t leaf_order;       /* (LBVH_LEAF_PRIMITIVE_ID) */
  uint32_t isect_mode;       /* LBVH_ISECT_* (REFERENCE) */
  uint32_t reserve_prims;    /* lanes' build workspaces sized at creation for
                                scenes of up to this many triangles, rounded
                                up to a power of two (0: none); a larger scene
                                grows a lane's workspace to the next power of
                                two, which waits for the device */
  uint32_t reserved[4];
} lbvh_server_opts;

/* lbvh_request.flags */
enum
{
  /* tri48, rays32 and the outputs are device memory: nothing is staged or
   * copied; a new scene borrows tri48 (as lbvh_build_triangles_device), a
   * refit borrows its new triangles in place of the old ones, and results
   * are in the output buffers when the request completes.  Borrowed memory
   * must stay valid while the scene is resident. */
  LBVH_REQUEST_DEVICE = 1
};

typedef struct lbvh_request
{
  uint64_t scene;      /* 0: a new scene from tri48; else a resident scene id */
  const float* tri48;  /* new scene: its triangles; resident scene: NULL, or
                          new triangle data of the same count (refit) */
  uint64_t n_tri;
  const float* rays32; /* rays */
  uint64_t n_rays;
  uint32_t* prim;      /* outputs, as lbvh_trace_closest */
  float* t;
  float* uv;           /* may be NULL */
  uint32_t flags;      /* LBVH_REQUEST_* */
  uint32_t reserved;
} lbvh_request;

typedef struct lbvh_request_info
{
  uint64_t scene;
  lbvh_status status;
  uint32_t coalesced_with; /* requests in the traversal launch it joined */
  double submit_us;        /* host clock, microseconds since server creation */
  double admit_us;
  double ready_us;         /* its build / refit / upload finished */
  double done_us;          /* results in the caller's buffers */
} lbvh_request_info;

typedef struct lbvh_server_stats
{
  uint64_t completed;
  uint64_t failed;
  uint64_t launches;           /* traversal launches */
  uint64_t coalesced_requests; /* requests summed over launches */
  uint64_t max_coalesced;
  uint64_t rays;
  uint64_t evictions;
  uint64_t resident_scenes;
  uint64_t memory_cap_bytes;
  uint64_t memory_used_bytes;
  uint64_t memory_peak_bytes;
} lbvh_server_stats;

void lbvh_server_opts_init(lbvh_server_opts* o);
lbvh_status lbvh_server_create(int device, const lbvh_server_opts* opts,
                               lbvh_server** out);
/* Finishes every submitted request, then frees everything. */
void lbvh_server_destroy(lbvh_server* server);
/* Queues a request; returns at once with its ticket and its scene id (the
 * new scene's id for a new scene). */
lbvh_status lbvh_server_submit(lbvh_server* server, const lbvh_request* req,
                               uint64_t* ticket, uint64_t* scene);
/* Waits for a request and collects it: its status is returned. */
lbvh_status lbvh_server_wait(lbvh_server* server, uint64_t ticket,
                             lbvh_request_info* info);
/* Collects a finished request (its status), or LBVH_E_BUSY while it runs. */
lbvh_status lbvh_server_poll(lbvh_server* server, uint64_t ticket,
                             lbvh_request_info* info);
/* Marks an idle resident scene as first to evict (LBVH_E_BUSY while
 * requests name it). */
lbvh_status lbvh_server_release_scene(lbvh_server* server, uint64_t scene);
lbvh_status lbvh_server_get_stats(lbvh_server* server, lbvh_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* LBVH_MI355X_H */
