// Host runtime internals shared by the C ABI translation units: the context
// and tree handles, the context's small-word layout, error helpers and the
// grow-only workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "engine.hpp"
#include "lbvh_mi355x.h"

namespace lbvh_rt {

using namespace lbvh_eng;

struct DevBuf
{
  void* p = nullptr;
  uint64_t cap = 0;
};

// One timed call: its kind and the events bracketing its phases.
struct EventSet
{
  int kind = 0; // 1 build, 2 trace, 3 refit
  int recorded = 0;
  hipEvent_t ev[8] = {};
};

// small workspace region layout (32-bit words)
constexpr uint32_t kSmallSpare = 0;         // unused
constexpr uint32_t kSmallErr = 1;          // cleared by the host after reading
constexpr uint32_t kSmallTickets = 2;      // kMaxSortPasses words
constexpr uint32_t kSmallTreeletCount = kSmallTickets + kMaxSortPasses;
constexpr uint32_t kSmallTreeletQueue = kSmallTreeletCount + 1;
constexpr uint32_t kSmallHist = kSmallTreeletQueue + 1;

// One timed candidate of the sort autotuner.
struct TuneRecord
{
  int batch;
  int key_bits;
  int log2n;
  SortConfig cfg;
  float ms; // device time of the whole build (or traversal) with this configuration
  bool picked;
  int scalar_bits = 32; // 32: builder<float> builds, 64: builder<double> builds
  int kind = 0;         // 0: a build's sort, 1: a traversal
  TraceConfig trace{};  // kind 1: the traversal configuration
  int trace_mode = 0;   // kind 1: the intersector mode
};

} // namespace lbvh_rt

struct lbvh_ctx
{
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int timing = 0; // 0 off, 1 per phase, 2 per call
  std::vector<lbvh_rt::EventSet> sets; // pool; sets[0..used) belong to timed calls
  size_t used = 0;
  std::string last_error;
  uint32_t epoch_next = 1;
  bool autotune = true;
  std::map<uint64_t, lbvh_eng::SortConfig> tuned; // by tune_key()
  std::map<uint64_t, lbvh_eng::TraceConfig> tuned_trace; // by trace_tune_key()
  std::vector<lbvh_rt::TuneRecord> tune_log;
  // workspace
  lbvh_rt::DevBuf cent, keys_a, keys_b, ids_a, ids_b, partials, small, status, counts,
    treelets, io_in, io_out, tree, tree_tickets;
  // traversal order of a ray batch: keys and ray indices (ping-pong), sort
  // chunk counts, histograms
  lbvh_rt::DevBuf ray_keys, ray_ids, ray_counts, ray_hist;
  // the treelet list in `treelets` belongs to the tree with this id (0: none);
  // its count and the fit's pull queue live in treelet_words
  uint64_t treelets_of = 0;
  lbvh_rt::DevBuf treelet_words;
};

// A built tree: node array, parent links, fit flags, centroid bounds, and
// the primitives (owned copy or borrowed pointer).
struct lbvh_bvh
{
  int device = 0;
  uint64_t n = 0;
  uint32_t code_bits = 30;
  uint32_t leaf_order = 0;
  int prim_kind = 0; // 0 triangles, 1 aabbs
  void* d_nodes = nullptr;
  uint32_t* d_parent = nullptr;
  uint32_t* d_flags = nullptr;
  float* d_bounds = nullptr;
  float* d_prims = nullptr;
  bool owns_prims = false;
  void* d_codes = nullptr;
  uint32_t* d_ids = nullptr;
  uint32_t* d_leafprim = nullptr; // primitive of each leaf position (primitive-id leaves)
  uint32_t* d_other_end = nullptr; // other end of each internal node's leaf range
  void* block = nullptr;           // the one allocation holding every buffer above
  uint64_t bytes = 0;
  lbvh_eng::SortConfig sort;       // the sort configuration the build used
  uint64_t id = 0;                 // unique per tree in the process (refit caches)
};

namespace lbvh_rt {

lbvh_status fail(lbvh_ctx* c, lbvh_status s, const std::string& msg);
// Device pointers the kernels read or write with vector loads and stores:
// triangles and rays need 16-byte alignment, boxes, uv pairs and 64-bit
// offsets 8, ids and distances 4.
inline bool
misaligned(const void* p, uintptr_t align)
{
  return (reinterpret_cast<uintptr_t>(p) & (align - 1)) != 0;
}
// Checks the device outputs of a traversal (prim, t, uv may be null).
inline bool
trace_outputs_misaligned(const void* prim, const void* t, const void* uv)
{
  return misaligned(prim, 4) || misaligned(t, 4) || misaligned(uv, 8);
}
// Rays per traversal call (and per coalesced server launch): a launch's
// work-item count must fit the dispatch's 32-bit grid size.
constexpr uint64_t kMaxRaysPerCall = (uint64_t(1) << 31) - 1;
lbvh_status hip_fail(lbvh_ctx* c, hipError_t e, const char* what);
lbvh_status ensure(lbvh_ctx* c, DevBuf& b, uint64_t bytes, bool zero_new = false);
void dev_free(void* p);
lbvh_status set_device(lbvh_ctx* c);
// Finds this device's copy of the allocation-free build's scratch array (the
// device must be current; once per device).
lbvh_status noalloc_init(lbvh_ctx* c);
void begin_timed(lbvh_ctx* c, int kind);
void record(lbvh_ctx* c, int idx, bool final = false);
bool valid_sort_config(const SortConfig& c, int key_bits);
lbvh_status sync_and_check(lbvh_ctx* c);

// ---- sort autotuner ----------------------------------------------------------
// The default configuration, the candidates timed by the tuner, and whether
// build options override the choice (then the tuner is not consulted).
SortConfig default_sort();
// The sort a single build uses without tuning (tuner off, or before a
// bucket is tuned): the hybrid from 65,536 keys, else the default.
SortConfig untuned_sort(int key_bits, uint64_t n);
// Grows a context's workspace to what a build of n primitives with this
// code width and sort configuration needs.
lbvh_status reserve_build(lbvh_ctx* c, uint64_t n, int code_bits, const SortConfig& cfg,
                          bool with_ids);
std::vector<SortConfig> sort_candidates(bool batch, int key_bits, uint64_t n);
bool sort_overridden(const lbvh_build_opts& o);
SortConfig sort_from_opts(const lbvh_build_opts& o, int key_bits, uint64_t n);
bool same_sort(const SortConfig& a, const SortConfig& b);
// Hybrid sort: the fewest top digits whose buckets hold at most
// kHybridMeanBucket keys on average (uniform keys), leaving the rest to the
// LDS sort.
int hybrid_global_digits(int key_bits, int digit_bits, uint64_t n);
constexpr uint64_t kHybridMeanBucket = 64;
// FP64 builds have buckets of their own: their pipeline around the sort
// differs (no encode-written chunk counts), so a pick timed on one kind of
// build does not carry over to the other.
uint64_t tune_key(bool batch, int key_bits, uint64_t n, int scalar_bits = 32);
int ceil_log2_u64(uint64_t x);
// Device milliseconds of the work `queue` puts on the context stream between
// its begin_timed()/record() calls (per-call timing forced on meanwhile).
template<class Q>
float time_call(lbvh_ctx* c, const Q& queue, lbvh_status* st);
// Sizes below this are launch-bound: the default is used without tuning.
constexpr uint64_t kTuneMinPrims = 1u << 16;

} // namespace lbvh_rt

#define HIP_TRY(ctx, expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return lbvh_rt::hip_fail((ctx), e_, #expr);                              \
  } while (0)

template<class Q>
float
lbvh_rt::time_call(lbvh_ctx* c, const Q& queue, lbvh_status* st)
{
  const int saved_timing = c->timing;
  const size_t saved_used = c->used;
  c->timing = 2;
  *st = queue();
  float ms = -1.0f;
  if (*st == LBVH_OK && c->used > saved_used) {
    EventSet& es = c->sets[c->used - 1];
    if (es.recorded >= 2 && hipEventSynchronize(es.ev[1]) == hipSuccess &&
        hipEventElapsedTime(&ms, es.ev[0], es.ev[1]) != hipSuccess)
      ms = -1.0f;
  }
  c->timing = saved_timing;
  c->used = saved_used;
  return ms;
}
