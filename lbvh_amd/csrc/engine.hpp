// Host-side launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbvh_eng {

constexpr uint32_t kMaxSortPasses = 16;

// Radix-sort launch configuration chosen by the tuner.
struct SortConfig
{
  int digit_bits = 8; // 8 or 10 (radix 256 / 1024)
  int items = 8;      // keys per thread; tile = 256 * items
  int strategy = 0;   // 0: onesweep (decoupled look-back), 1: reduce-then-scan,
                      // 2: hybrid (top digits by reduce-then-scan, then every
                      //    run of equal top bits sorted by the full key in LDS)
  int groups = 0;     // reduce-then-scan: target number of chunks (0 = about two tiles each, 512-2,048)
  int global_digits = 0; // hybrid: digits sorted by global passes (the top ones)
};

// Hybrid strategy: keys per workgroup of the LDS sort (nominal and maximum).
constexpr uint32_t kHybridTileKeys = 2048;
constexpr uint32_t kHybridTileCap = 4096;

inline int
sort_passes(int key_bits, int digit_bits)
{
  return (key_bits + digit_bits - 1) / digit_bits;
}

// Scratch needed by the sort for n keys.
struct SortWorkspace
{
  uint64_t* status = nullptr;  // [nblocks * radix] look-back words
  uint64_t status_words = 0;
  uint32_t* tickets = nullptr; // [kMaxSortPasses] virtual block counters
  uint32_t* hist = nullptr;    // [passes * radix] global digit histograms
  uint32_t* err = nullptr;     // bounded-wait failure flag
  uint32_t* counts = nullptr;  // reduce-then-scan chunk counts [radix * groups]
  uint64_t counts_words = 0;
  bool first_counts = false;   // the first reduce-then-scan pass's counts are in place
  // Scene batches (keys (scene << 30 | code), grouped by scene in ascending
  // order): chunks that each lie in one scene.  When set and the hybrid's
  // global digits hold at most one digit of code bits, one pass over these
  // chunks (counts chunk-major, [groups * radix]) replaces the global passes.
  const uint32_t* seg_chunks = nullptr; // [2 * seg_groups]: begin, end of each chunk
  const uint32_t* seg_first = nullptr;  // [nseg + 1]: first chunk of each scene
  const uint64_t* seg_off = nullptr;    // [nseg + 1]: scene offsets
  uint32_t seg_groups = 0;
  uint32_t nseg = 0;
};

// True when the hybrid sort of batch keys (code bits below bit 30) runs as
// one scene-segmented pass plus the LDS sort (see SortWorkspace::seg_chunks).
bool seg_pass_applies(int key_bits, const SortConfig& cfg);
// Keys per scene-aligned chunk of that pass (a multiple of every sort tile).
uint64_t seg_chunk_keys(uint64_t n);

// centroid of every primitive box (SoA) and one partial box of centroid
// bounds per workgroup (centroid_partials_needed(n) boxes of 6 floats); also
// zeroes `zero_words` 32-bit words at `zero_ptr` (histograms, tickets, error
// flag).
hipError_t launch_centroid_bounds(const float* prims, int kind, uint64_t n, float* cx,
                                  float* cy, float* cz, float* partials, uint32_t* zero_ptr,
                                  uint32_t zero_words, hipStream_t s);
uint32_t centroid_partials_needed(uint64_t n);

// Morton codes (u32 for 30-bit, u64 for 63-bit) and the digit histograms of
// passes first_pass .. passes - 1, added into hist[p << digit_bits ...] (zero
// before).  The centroid bounds are the union of launch_centroid_bounds'
// partials for the same n; they are also written to `bounds` (6 floats).
// counts != nullptr: `groups` workgroups encode chunks of `chunk` keys and
// also write the chunk counts of pass first_pass (counts[d * groups + g]),
// the reduce-then-scan layout.
hipError_t launch_encode(int code_bits, const float* cx, const float* cy, const float* cz,
                         uint64_t n, const float* partials, float* bounds, void* codes,
                         uint32_t* hist,
                         int digit_bits, int first_pass, int passes, uint32_t* counts,
                         uint32_t groups, uint64_t chunk, hipStream_t s);
uint32_t encode_groups(uint64_t n);

// LSD onesweep radix sort of (key, id) pairs in place across ping-pong
// buffers; ids_in == nullptr means ids = 0..n-1; ids_a == nullptr sorts keys
// only.  Returns in *result_in_b whether the sorted data ended in the B
// buffers.  `epoch` must change between sorts sharing `ws.status`.
// key_bytes: 4 (u32 keys) or 8 (u64 keys); key_bits: the low bits that vary.
hipError_t launch_sort(int key_bits, int key_bytes, uint64_t n, void* keys_a, void* keys_b,
                       uint32_t* ids_a,
                       uint32_t* ids_b, const SortConfig& cfg, const SortWorkspace& ws,
                       uint32_t epoch, bool* result_in_b, hipStream_t s);
uint64_t sort_status_words(uint64_t n, const SortConfig& cfg);
// 32-bit words of ws.counts the sort needs for n keys.
uint64_t sort_counts_words(uint64_t n, const SortConfig& cfg);
// Number of chunks of the reduce-then-scan strategy (and tiles per chunk).
uint32_t rts_groups(uint64_t n, const SortConfig& cfg, uint32_t* chunk_tiles);
// Keys per chunk of the reduce-then-scan strategy.
uint64_t rts_chunk_keys(uint64_t n, const SortConfig& cfg);

// Scene batches: prim_off == nullptr is one scene of n primitives; otherwise
// scene s holds primitives [prim_off[s], prim_off[s + 1]) (n = prim_off[nseg],
// every scene at least 2), its internal nodes start at prim_off[s] - s and its
// parent links at 2 prim_off[s] - s.  Batched sorted keys are
// (scene << 30 | 30-bit code) in u64.

// Parent links (refits): internal node c at parent[c], the leaf at sorted
// position k at parent[n - 1 + k], per scene; bit 31: the parent is a top
// node (more than 1,024 leaves), bit 30: right child, bit 29: the sibling is
// a leaf; the parent's scene-local index below.

// Bottom-up box fit of an existing tree (refits): treelets from the list in
// LDS, then the top tree with device-scope flags (flags must be zero and are
// left zero; *queue must be zero).
// leafprim == nullptr: the leaf at position k is primitive k of its scene.
hipError_t launch_fit(const float* prims, int kind, uint64_t n, const uint64_t* prim_off,
                      uint32_t nseg, void* nodes, const uint32_t* parent,
                      const uint32_t* other_end, const uint32_t* leafprim, uint32_t* flags,
                      const void* treelets, uint32_t* treelet_count, uint32_t* queue,
                      hipStream_t s);

// Rebuilds the treelet list of a built tree (for a refit); *treelet_count
// must be zero.
// Bottom-up tree construction and fit in one kernel (levels of workgroups
// merging in LDS).  Writes the node records, parent links,
// other_end, leafprim (ids != nullptr and primitive-id leaves) and leaves
// the fit flags at 0.  `scratch` holds tree_scratch_bytes(n) bytes;
// `tickets` holds tree_ticket_words(n) words that were zero when first used
// (the kernel leaves them at zero).
hipError_t launch_tree(int code_bits, const void* sorted, const uint32_t* ids, uint64_t n,
                       const uint64_t* prim_off, uint32_t nseg, const float* prims, int kind,
                       void* nodes, uint32_t* parent, uint32_t* leafprim, uint32_t* flags,
                       uint32_t* other_end, void* scratch, uint32_t* tickets, hipStream_t s);
uint64_t tree_scratch_bytes(uint64_t n);
uint64_t tree_ticket_words(uint64_t n);
hipError_t launch_enum_treelets(uint64_t n, const uint64_t* prim_off, uint32_t nseg,
                                const uint32_t* parent, const uint32_t* other_end,
                                void* treelets, uint32_t* treelet_count, hipStream_t s);

// Closest-hit traversal over the node records (lbvh::node<float>).  For a
// batch (prim_off != nullptr) ray r belongs to the scene s with ray_off[s] <=
// r < ray_off[s + 1] and hits report scene-local primitive ids.  order (one
// tree only, may be null): a permutation of the rays; the launch takes them
// in that order and writes each hit to the ray's own slot.  waves_per_simd
// (0 or 8: as many as the registers allow) caps the waves sharing a SIMD.
hipError_t launch_trace(int mode, const void* nodes, const float* tris, uint64_t n,
                        const uint64_t* prim_off, const uint64_t* ray_off, uint32_t nseg,
                        const float* rays, uint64_t nrays, uint32_t* prim, float* t, float* uv,
                        hipStream_t s, const uint32_t* order = nullptr, int waves_per_simd = 0);

// Traversal configuration chosen by the tuner: whether rays are put in the
// order of their keys first, and the waves per SIMD.
struct TraceConfig
{
  int reorder = 0;
  int waves_per_simd = 8;
};

// Sort keys of a ray batch for the traversal order (kRayKeyBits bits: the
// direction octant above a coarse Morton code of a point along the ray, over
// the tree's centroid bounds): `groups` workgroups of `chunk` rays each
// write their keys, the first 8-bit pass's chunk counts (counts[d * groups +
// g]) and add both passes' digit counts to hist[0 .. 511] (zero before).
constexpr int kRayKeyBits = 15;
hipError_t launch_ray_keys(const float* rays, uint64_t nrays, const float* bounds,
                           uint32_t* keys, uint32_t* hist, uint32_t* counts, uint32_t groups,
                           uint64_t chunk, hipStream_t s);

// One tree's share of a coalesced traversal launch: its rays are global rays
// [ray_begin, next job's ray_begin) (the last job ends at the launch's ray
// count) and read from / written to the job's own buffers.
struct TraceJob
{
  const void* nodes; // node records
  const float* tris;
  const float* rays;
  uint32_t* prim;
  float* t;
  float* uv; // may be null
  uint64_t ray_begin;
  uint32_t nprims;
  uint32_t pad;
};

// Closest hit over many independent trees in one launch; `jobs` is device
// memory, sorted by ray_begin with jobs[0].ray_begin = 0.
hipError_t launch_trace_jobs(int mode, const TraceJob* jobs, uint32_t njobs, uint64_t nrays,
                             hipStream_t s);

// Batched centroid bounds: the centre (SoA) of every primitive and per-scene
// centroid bounds as order-preserving integers (6 per scene: min xyz, max
// xyz; launch_scene_bounds_init resets them first).
hipError_t launch_scene_bounds_init(int32_t* bounds_key, uint32_t nseg, hipStream_t s);
hipError_t launch_scene_bounds(const float* prims, int kind, uint64_t n, const uint64_t* prim_off,
                               uint32_t nseg, float* cx, float* cy, float* cz, int32_t* bounds_key,
                               hipStream_t s);
// Batched 30-bit codes keyed (scene << 30 | code) and the digit histograms
// of radix passes first_pass .. passes - 1 (per-workgroup rows in part_hist,
// summed into hist[p << digit_bits ...], which must be zero).  counts !=
// nullptr: `groups` workgroups encode chunks of `chunk` keys and write the
// first pass's chunk counts (counts[d * groups + g]) as launch_encode does.
// chunks != nullptr as well: workgroup g encodes [chunks[2g], chunks[2g + 1])
// and writes its counts chunk-major (counts[g * radix + d]); part_hist and
// hist are then unused (may be nullptr) and passes must be first_pass + 1.
hipError_t launch_scene_encode(const float* cx, const float* cy, const float* cz, uint64_t n,
                               const uint64_t* prim_off, uint32_t nseg, const int32_t* bounds_key,
                               uint64_t* keys, uint32_t* part_hist, uint32_t* hist,
                               int digit_bits, int first_pass, int passes, uint32_t* counts,
                               uint32_t groups, uint64_t chunk, const uint32_t* chunks,
                               hipStream_t s);
uint32_t scene_encode_groups(uint64_t n);

// builder<double>: centroid bounds of FP64 boxes (6 doubles each, 16-byte
// aligned) via f64_bounds_blocks(n) partials of 6 doubles; u64 codes with the
// digit histograms of passes first_pass .. passes - 1 added to hist[p <<
// digit_bits ...] (zero before); and the tree over sorted
// codes into node<double> records (64 B) with f64_tree_scratch_bytes(n) of
// scratch (parent links, range ends, fit flags cleared by the launch,
// staging boxes).  ids != nullptr: leaves name ids[position] instead of the
// position.
uint32_t f64_bounds_blocks(uint64_t n);
hipError_t launch_f64_bounds(const double* boxes, uint64_t n, double* partials, double* bounds,
                             hipStream_t s);
hipError_t launch_f64_encode(const double* boxes, uint64_t n, const double* bounds,
                             uint64_t* codes, uint32_t* hist, int digit_bits, int first_pass,
                             int passes, hipStream_t s);
uint64_t f64_tree_scratch_bytes(uint64_t n);
hipError_t launch_f64_tree(const uint64_t* sorted, const uint32_t* ids, const double* boxes,
                           uint64_t n, void* nodes, void* scratch, hipStream_t s);

// Closest-hit traversal of a node<double> array (n - 1 records of 64 B for
// n primitives) over triangle<double> records (12 doubles) for ray<double>
// records (8 doubles); mode 0 = reference, 1 = lowest id on equal distance.
// prim: u64 (UINT64_MAX = miss), t: f64, uv: 2 f64 per ray or nullptr.
hipError_t launch_trace_f64(int mode, const void* nodes, const double* tris, uint64_t n,
                            const double* rays, uint64_t nrays, uint64_t* prim, double* t,
                            double* uv, hipStream_t s);

// Fill `words` 32-bit words with `value`.
hipError_t launch_fill_u32(uint32_t* p, uint64_t words, uint32_t value, hipStream_t s);

} // namespace lbvh_eng
