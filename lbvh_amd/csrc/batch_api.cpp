// C ABI of scene batches (include/lbvh_mi355x.h, "scene batches").
//
// A batch holds many triangle scenes in one array and builds them in one
// segmented pass.  It owns every device buffer its rebuilds, refits and
// traversals need (scratch included), so a rebuild or refit allocates
// nothing, does not synchronize, and can be captured once into a HIP graph
// and replayed.  Clearing inside a captured sequence uses fill kernels only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"
#include "lbvh_mi355x.h"
#include "runtime.hpp"

using namespace lbvh_eng;
using namespace lbvh_rt;

namespace {

constexpr uint64_t kCodeMask = 0x3fffffffull;

struct Graph
{
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr; // the stream it was captured on
  const float* prims = nullptr; // the primitive pointer it reads
  void reset()
  {
    if (exec)
      (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    stream = nullptr;
    prims = nullptr;
  }
};

} // namespace

struct lbvh_batch
{
  int device = 0;
  uint64_t n = 0;
  uint32_t nseg = 0;
  uint32_t leaf_order = 0;
  bool with_ids = false;
  int key_bits = 30;
  int passes = 0;
  SortConfig cfg;
  std::vector<uint64_t> h_off; // [nseg + 1]
  float* d_prims = nullptr;
  bool owns_prims = false;
  uint64_t* d_off = nullptr;
  // tree
  void* d_nodes = nullptr;
  uint32_t* d_parent = nullptr;
  uint32_t* d_flags = nullptr;
  uint32_t* d_other_end = nullptr;
  uint32_t* d_leafprim = nullptr;
  int32_t* d_bounds_key = nullptr;
  // scratch
  float* d_cent = nullptr;
  uint64_t* d_keys_a = nullptr;
  uint64_t* d_keys_b = nullptr;
  uint32_t* d_ids_a = nullptr;
  uint32_t* d_ids_b = nullptr;
  uint32_t* d_part_hist = nullptr;
  uint32_t* d_small = nullptr;
  uint32_t small_words = 0;
  uint32_t* d_counts = nullptr;
  uint64_t counts_words = 0;
  // scene-aligned sort chunks: [2 * seg_groups] begin/end, then [nseg + 1]
  // first chunk of each scene
  uint32_t* d_seg = nullptr;
  uint32_t seg_groups = 0;
  void* d_treelets = nullptr;
  void* d_tree = nullptr;           // tree kernel segments
  uint32_t* d_tree_tickets = nullptr; // tree kernel tickets (zero at rest)
  bool sorted_in_b = false;
  uint64_t bytes = 0;
  Graph g_build, g_refit;
};

namespace {

void
free_batch(lbvh_batch* b)
{
  if (!b)
    return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(b->device);
  b->g_build.reset();
  b->g_refit.reset();
  void* bufs[] = { b->d_off,       b->d_nodes,  b->d_parent,      b->d_flags,   b->d_other_end,
                   b->d_leafprim,  b->d_bounds_key,  b->d_cent,    b->d_keys_a,
                   b->d_keys_b,    b->d_ids_a,  b->d_ids_b,       b->d_part_hist, b->d_small,
                   b->d_counts,    b->d_treelets, b->d_tree, b->d_tree_tickets, b->d_seg };
  for (void* p : bufs)
    dev_free(p);
  if (b->owns_prims)
    dev_free(b->d_prims);
  (void)hipSetDevice(cur);
  delete b;
}

int
ceil_log2(uint32_t x)
{
  int b = 0;
  while ((uint64_t(1) << b) < x)
    b++;
  return b;
}

// Chunks of at most seg_chunk_keys(n) keys that each lie in one scene, in
// key order: begin/end pairs, then the first chunk of every scene (nseg + 1).
std::vector<uint32_t>
seg_table(const std::vector<uint64_t>& off, uint64_t n, uint32_t* groups)
{
  const uint64_t ck = seg_chunk_keys(n);
  const size_t nseg = off.size() - 1;
  std::vector<uint32_t> pairs, first(nseg + 1);
  for (size_t s = 0; s < nseg; s++) {
    first[s] = uint32_t(pairs.size() / 2);
    for (uint64_t b = off[s]; b < off[s + 1]; b += ck) {
      pairs.push_back(uint32_t(b));
      pairs.push_back(uint32_t(std::min(off[s + 1], b + ck)));
    }
  }
  first[nseg] = uint32_t(pairs.size() / 2);
  *groups = first[nseg];
  pairs.insert(pairs.end(), first.begin(), first.end());
  return pairs;
}

// Scratch is sized for every sort configuration in `cfgs` (the tuner's
// candidates, or the one configuration in use).
lbvh_status
alloc_batch(lbvh_ctx* c, lbvh_batch* b, const std::vector<SortConfig>& cfgs)
{
  const uint64_t n = b->n;
  const uint64_t m = n - b->nseg; // internal nodes
  uint64_t hist_words = 0;
  uint64_t hist_rows = scene_encode_groups(n);
  b->small_words = 0;
  b->counts_words = 0;
  uint32_t seg_groups = 0;
  const std::vector<uint32_t> seg = seg_table(b->h_off, n, &seg_groups);
  bool any_seg = false;
  for (const SortConfig& cfg : cfgs) {
    if (seg_pass_applies(b->key_bits, cfg)) {
      any_seg = true;
      const uint64_t cw =
        (uint64_t(seg_groups) << cfg.digit_bits) + 2 + 2 * (n / kHybridTileCap + 1);
      if (cw > b->counts_words)
        b->counts_words = cw;
    }
    // reduce-then-scan and hybrid: one encode workgroup per sort chunk
    hist_rows = std::max<uint64_t>(hist_rows, rts_groups(n, cfg, nullptr));
    const uint32_t radix = 1u << cfg.digit_bits;
    const uint32_t words = uint32_t(sort_passes(b->key_bits, cfg.digit_bits)) * radix;
    if (words > hist_words)
      hist_words = words;
    if (kSmallHist + words > b->small_words)
      b->small_words = kSmallHist + words;
    const uint64_t cw = sort_counts_words(n, cfg);
    if (cw > b->counts_words)
      b->counts_words = cw;
  }
  struct
  {
    void** p;
    uint64_t bytes;
  } req[] = {
    { reinterpret_cast<void**>(&b->d_off), 8ull * (b->nseg + 1) },
    { &b->d_nodes, m * 32 },
    { reinterpret_cast<void**>(&b->d_parent), (2 * n - b->nseg) * 4 },
    { reinterpret_cast<void**>(&b->d_flags), m * 4 },
    { reinterpret_cast<void**>(&b->d_other_end), m * 4 },
    { reinterpret_cast<void**>(&b->d_leafprim),
      b->leaf_order == LBVH_LEAF_PRIMITIVE_ID ? n * 4 : 0 },
    { reinterpret_cast<void**>(&b->d_bounds_key), 24ull * b->nseg },
    { reinterpret_cast<void**>(&b->d_cent), 12 * n },
    { reinterpret_cast<void**>(&b->d_keys_a), 8 * n },
    { reinterpret_cast<void**>(&b->d_keys_b), 8 * n },
    { reinterpret_cast<void**>(&b->d_ids_a), b->with_ids ? 4 * n : 0 },
    { reinterpret_cast<void**>(&b->d_ids_b), b->with_ids ? 4 * n : 0 },
    { reinterpret_cast<void**>(&b->d_part_hist), 4ull * hist_rows * hist_words },
    { reinterpret_cast<void**>(&b->d_small), 4ull * b->small_words },
    { reinterpret_cast<void**>(&b->d_counts), 4 * b->counts_words },
    { &b->d_treelets, 8 * n },
    { &b->d_tree, tree_scratch_bytes(n) },
    { reinterpret_cast<void**>(&b->d_tree_tickets), 4 * tree_ticket_words(n) },
    { reinterpret_cast<void**>(&b->d_seg), any_seg ? 4 * seg.size() : 0 },
  };
  for (auto& r : req) {
    if (!r.bytes)
      continue;
    hipError_t e = hipMalloc(r.p, r.bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, LBVH_E_OOM, "device allocation for the batch failed");
    }
    b->bytes += r.bytes;
  }
  hipError_t e = hipMemset(b->d_tree_tickets, 0, 4 * tree_ticket_words(n));
  if (e != hipSuccess)
    return hip_fail(c, e, "clearing the tree tickets");
  if (any_seg) {
    e = hipMemcpy(b->d_seg, seg.data(), 4 * seg.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return hip_fail(c, e, "the scene chunk table");
    b->seg_groups = seg_groups;
  }
  return LBVH_OK;
}

// Queues a full rebuild of every scene (no allocation, no host sync).
hipError_t
queue_build(lbvh_ctx* c, lbvh_batch* b, bool timed)
{
  const uint64_t n = b->n;
  hipStream_t s = c->stream;
  uint32_t* small = b->d_small;
  SortWorkspace ws;
  ws.status = nullptr;
  ws.status_words = 0;
  ws.tickets = small + kSmallTickets;
  ws.hist = small + kSmallHist;
  ws.err = small + kSmallErr;
  ws.counts = b->d_counts;
  ws.counts_words = b->counts_words;
  // reduce-then-scan and the hybrid (what batches sort with): the encode
  // writes the first pass's chunk counts
  const bool chunked = b->cfg.strategy == 1 || b->cfg.strategy == 2;
  ws.first_counts = chunked;
  const int first_pass = b->cfg.strategy == 2 ? b->passes - b->cfg.global_digits : 0;
  // the hybrid of one code digit: one pass over scene-aligned chunks
  const bool seg = b->d_seg && seg_pass_applies(b->key_bits, b->cfg);
  if (seg) {
    ws.seg_chunks = b->d_seg;
    ws.seg_first = b->d_seg + 2 * size_t(b->seg_groups);
    ws.seg_off = b->d_off;
    ws.seg_groups = b->seg_groups;
    ws.nseg = b->nseg;
  }
  if (timed)
    record(c, 0);
  hipError_t e = launch_fill_u32(small + kSmallTickets, b->small_words - kSmallTickets, 0u, s);
  if (e == hipSuccess)
    e = launch_scene_bounds_init(b->d_bounds_key, b->nseg, s);
  if (e == hipSuccess)
    e = launch_scene_bounds(b->d_prims, 0, n, b->d_off, b->nseg, b->d_cent, b->d_cent + n,
                            b->d_cent + 2 * n, b->d_bounds_key, s);
  if (e == hipSuccess)
    e = launch_scene_encode(b->d_cent, b->d_cent + n, b->d_cent + 2 * n, n, b->d_off, b->nseg,
                            b->d_bounds_key, b->d_keys_a, seg ? nullptr : b->d_part_hist,
                            seg ? nullptr : ws.hist, b->cfg.digit_bits, first_pass,
                            seg ? first_pass + 1 : b->passes, chunked ? b->d_counts : nullptr,
                            seg ? b->seg_groups : chunked ? rts_groups(n, b->cfg, nullptr) : 0u,
                            chunked && !seg ? rts_chunk_keys(n, b->cfg) : 0,
                            seg ? b->d_seg : nullptr, s);
  if (timed)
    record(c, 1);
  bool in_b = false;
  if (e == hipSuccess)
    e = launch_sort(b->key_bits, 8, n, b->d_keys_a, b->d_keys_b, b->with_ids ? b->d_ids_a : nullptr,
                    b->with_ids ? b->d_ids_b : nullptr, b->cfg, ws, 0u, &in_b, s);
  b->sorted_in_b = in_b;
  if (timed)
    record(c, 2);
  const uint64_t* sorted = in_b ? b->d_keys_b : b->d_keys_a;
  const uint32_t* sorted_ids = b->with_ids ? (in_b ? b->d_ids_b : b->d_ids_a) : nullptr;
  // tree and fit in one kernel (the fit phase is part of the tree phase)
  const uint32_t* leaf_ids = b->leaf_order == LBVH_LEAF_PRIMITIVE_ID ? sorted_ids : nullptr;
  if (e == hipSuccess)
    e = launch_tree(30, sorted, leaf_ids, n, b->d_off, b->nseg, b->d_prims, 0, b->d_nodes,
                    b->d_parent, b->d_leafprim, b->d_flags, b->d_other_end, b->d_tree,
                    b->d_tree_tickets, s);
  if (timed) {
    record(c, 3);
    record(c, 4, true);
  }
  return e;
}

hipError_t
queue_refit(lbvh_ctx* c, lbvh_batch* b)
{
  hipStream_t s = c->stream;
  uint32_t* small = b->d_small;
  hipError_t e = launch_fill_u32(small + kSmallTreeletCount, 2, 0u, s);
  if (e == hipSuccess)
    e = launch_enum_treelets(b->n, b->d_off, b->nseg, b->d_parent, b->d_other_end,
                             b->d_treelets, small + kSmallTreeletCount, s);
  if (e == hipSuccess)
    e = launch_fit(b->d_prims, 0, b->n, b->d_off, b->nseg, b->d_nodes, b->d_parent,
                   b->d_other_end, b->d_leafprim, b->d_flags, b->d_treelets,
                   small + kSmallTreeletCount, small + kSmallTreeletQueue, s);
  return e;
}

// Runs `queue` directly, or through a graph captured from it on first use
// (recaptured when the stream or the primitive pointer changed).
template<class Q>
lbvh_status
run_maybe_graph(lbvh_ctx* c, lbvh_batch* b, Graph& g, bool use_graph, int kind, const Q& queue)
{
  if (!use_graph) {
    begin_timed(c, kind);
    hipError_t e = queue(true);
    if (e != hipSuccess)
      return hip_fail(c, e, "batch launch");
    return LBVH_OK;
  }
  if (g.exec && (g.stream != c->stream || g.prims != b->d_prims))
    g.reset();
  if (!g.exec) {
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = queue(false);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph)
        (void)hipGraphDestroy(graph);
      return hip_fail(c, e != hipSuccess ? e : e2, "batch graph capture");
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
      return hip_fail(c, e, "hipGraphInstantiate");
    g.exec = exec;
    g.stream = c->stream;
    g.prims = b->d_prims;
  }
  begin_timed(c, kind);
  record(c, 0);
  HIP_TRY(c, hipGraphLaunch(g.exec, c->stream));
  record(c, 1, true);
  return LBVH_OK;
}

lbvh_status
build_batch_common(lbvh_ctx* c, const float* d_tri, bool owns, const uint64_t* off,
                   uint32_t nseg, const lbvh_build_opts* opts_in, lbvh_batch** out)
{
  lbvh_build_opts opts;
  lbvh_build_opts_init(&opts);
  if (opts_in)
    opts = *opts_in;
  if (opts.code_bits != 30)
    return fail(c, LBVH_E_INVALID, "batches use 30-bit codes");
  if (opts.leaf_order > 1)
    return fail(c, LBVH_E_INVALID, "unknown leaf order");
  const uint64_t n = off[nseg];
  if (n >= (uint64_t(1) << 31))
    return fail(c, LBVH_E_INVALID, "a batch holds fewer than 2^31 triangles");
  if (opts.sort_strategy == 1) // its look-back state is tied to one build: no graph replays
    return fail(c, LBVH_E_INVALID, "batches do not use the onesweep sort");
  const int key_bits = 30 + ceil_log2(nseg);
  // explicit options, else the tuned choice for this key width and size,
  // else every candidate (timed below on this batch)
  std::vector<SortConfig> cands;
  bool tuning = false;
  if (sort_overridden(opts)) {
    const SortConfig cfg = sort_from_opts(opts, key_bits, n);
    if (!valid_sort_config(cfg, key_bits))
      return fail(c, LBVH_E_INVALID, "unsupported sort digit width / tile");
    cands.push_back(cfg);
  } else if (!c->autotune || n < kTuneMinPrims) {
    cands.push_back(default_sort());
  } else {
    auto it = c->tuned.find(tune_key(true, key_bits, n));
    if (it != c->tuned.end()) {
      cands.push_back(it->second);
    } else {
      cands = sort_candidates(true, key_bits, n);
      tuning = true;
    }
  }
  auto* b = new (std::nothrow) lbvh_batch;
  if (!b)
    return fail(c, LBVH_E_OOM, "host allocation failed");
  b->device = c->device;
  b->n = n;
  b->nseg = nseg;
  b->leaf_order = opts.leaf_order;
  b->with_ids = opts.leaf_order == LBVH_LEAF_PRIMITIVE_ID || opts.keep_codes;
  b->key_bits = key_bits;
  b->cfg = cands[0];
  b->passes = sort_passes(b->key_bits, b->cfg.digit_bits);
  b->h_off.assign(off, off + nseg + 1);
  b->d_prims = const_cast<float*>(d_tri);
  b->owns_prims = owns;
  if (owns)
    b->bytes += n * 48;
  lbvh_status st = alloc_batch(c, b, cands);
  if (st != LBVH_OK) {
    b->owns_prims = false;
    free_batch(b);
    return st;
  }
  hipError_t e = hipMemcpyAsync(b->d_off, off, 8ull * (nseg + 1), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream); // `off` may be pageable caller memory
  if (e != hipSuccess) {
    b->owns_prims = false;
    free_batch(b);
    return hip_fail(c, e, "batch offsets");
  }
  if (tuning) {
    // every candidate twice on this batch; the second run is timed
    float best = -1.0f;
    const size_t first = c->tune_log.size();
    SortConfig pick = cands[0];
    for (const SortConfig& cand : cands) {
      b->cfg = cand;
      b->passes = sort_passes(b->key_bits, cand.digit_bits);
      float ms = -1.0f;
      for (int rep = 0; rep < 2 && st == LBVH_OK; rep++) {
        ms = time_call(c, [&] {
          begin_timed(c, 1);
          hipError_t qe = queue_build(c, b, true);
          return qe == hipSuccess ? LBVH_OK : hip_fail(c, qe, "batch build");
        }, &st);
        if (st == LBVH_OK)
          st = sync_and_check(c);
      }
      if (st != LBVH_OK) {
        b->owns_prims = false;
        free_batch(b);
        return st;
      }
      c->tune_log.push_back(TuneRecord{ 1, b->key_bits, ceil_log2_u64(n), cand, ms, false });
      if (ms >= 0.0f && (best < 0.0f || ms < best)) {
        best = ms;
        pick = cand;
      }
    }
    for (size_t i = first; i < c->tune_log.size(); i++) {
      c->tune_log[i].picked = same_sort(c->tune_log[i].cfg, pick);
    }
    c->tuned[tune_key(true, b->key_bits, n)] = pick;
    b->cfg = pick;
    b->passes = sort_passes(b->key_bits, pick.digit_bits);
  }
  begin_timed(c, 1);
  e = queue_build(c, b, true);
  if (e != hipSuccess) {
    b->owns_prims = false;
    free_batch(b);
    return hip_fail(c, e, "batch build");
  }
  *out = b;
  return LBVH_OK;
}

lbvh_status
check_offsets(lbvh_ctx* c, const uint64_t* off, uint32_t nseg)
{
  if (!off || nseg == 0)
    return fail(c, LBVH_E_INVALID, "a batch needs scene offsets and at least one scene");
  if (off[0] != 0)
    return fail(c, LBVH_E_INVALID, "scene offsets start at 0");
  for (uint32_t s = 0; s < nseg; s++)
    if (off[s + 1] < off[s] + 2)
      return fail(c, LBVH_E_INVALID,
                  "scene " + std::to_string(s) + " has fewer than 2 triangles");
  return LBVH_OK;
}

} // namespace

extern "C" {

lbvh_status
lbvh_build_batch_device(lbvh_ctx* c, const float* d_tri, const uint64_t* prim_off,
                        uint32_t n_scenes, const lbvh_build_opts* opts, lbvh_batch** out)
{
  if (!c || !out || !d_tri)
    return fail(c, LBVH_E_INVALID, "null argument");
  *out = nullptr;
  if (misaligned(d_tri, 16))
    return fail(c, LBVH_E_INVALID, "device triangles must be 16-byte aligned");
  lbvh_status st = check_offsets(c, prim_off, n_scenes);
  if (st != LBVH_OK)
    return st;
  if ((st = set_device(c)) != LBVH_OK)
    return st;
  return build_batch_common(c, d_tri, false, prim_off, n_scenes, opts, out);
}

lbvh_status
lbvh_build_batch(lbvh_ctx* c, const float* tri, const uint64_t* prim_off, uint32_t n_scenes,
                 const lbvh_build_opts* opts, lbvh_batch** out)
{
  if (!c || !out || !tri)
    return fail(c, LBVH_E_INVALID, "null argument");
  *out = nullptr;
  lbvh_status st = check_offsets(c, prim_off, n_scenes);
  if (st != LBVH_OK)
    return st;
  if ((st = set_device(c)) != LBVH_OK)
    return st;
  const uint64_t bytes = prim_off[n_scenes] * 48;
  float* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, LBVH_E_OOM, "device allocation for triangles failed");
  }
  hipError_t e = hipMemcpyAsync(d, tri, bytes, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) {
    dev_free(d);
    return hip_fail(c, e, "hipMemcpyAsync");
  }
  st = build_batch_common(c, d, true, prim_off, n_scenes, opts, out);
  if (st != LBVH_OK) {
    (void)hipStreamSynchronize(c->stream);
    dev_free(d);
    return st;
  }
  st = sync_and_check(c);
  if (st != LBVH_OK) {
    free_batch(*out);
    *out = nullptr;
  }
  return st;
}

// Replaces the triangles a batch reads: owned storage is overwritten from
// device memory; a borrowed pointer is swapped.
static lbvh_status
update_prims(lbvh_ctx* c, lbvh_batch* b, const float* d_tri)
{
  if (!d_tri || d_tri == b->d_prims)
    return LBVH_OK;
  if (misaligned(d_tri, 16))
    return fail(c, LBVH_E_INVALID, "device triangles must be 16-byte aligned");
  if (b->owns_prims) {
    HIP_TRY(c, hipMemcpyAsync(b->d_prims, d_tri, b->n * 48, hipMemcpyDeviceToDevice, c->stream));
  } else {
    b->d_prims = const_cast<float*>(d_tri);
  }
  return LBVH_OK;
}

lbvh_status
lbvh_batch_rebuild(lbvh_ctx* c, lbvh_batch* b, const float* d_tri, int use_graph)
{
  if (!c || !b)
    return fail(c, LBVH_E_INVALID, "null argument");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if ((st = update_prims(c, b, d_tri)) != LBVH_OK)
    return st;
  return run_maybe_graph(c, b, b->g_build, use_graph != 0, 1,
                         [&](bool timed) { return queue_build(c, b, timed); });
}

lbvh_status
lbvh_batch_refit(lbvh_ctx* c, lbvh_batch* b, const float* d_tri, int use_graph)
{
  if (!c || !b)
    return fail(c, LBVH_E_INVALID, "null argument");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if ((st = update_prims(c, b, d_tri)) != LBVH_OK)
    return st;
  return run_maybe_graph(c, b, b->g_refit, use_graph != 0, 3, [&](bool timed) {
    if (timed)
      record(c, 0);
    hipError_t e = queue_refit(c, b);
    if (timed)
      record(c, 1, true);
    return e;
  });
}

lbvh_status
lbvh_batch_get_info(const lbvh_batch* b, lbvh_batch_info* out)
{
  if (!b || !out)
    return LBVH_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  out->n_scenes = b->nseg;
  out->leaf_order = b->leaf_order;
  out->n_prims = b->n;
  out->n_nodes = b->n - b->nseg;
  out->sort_digit_bits = uint32_t(b->cfg.digit_bits);
  out->sort_tile = uint32_t(b->cfg.items * 256);
  out->sort_passes = uint32_t(b->passes);
  out->sort_strategy = uint32_t(b->cfg.strategy + 1);
  out->device_bytes = b->bytes;
  out->d_nodes = b->d_nodes;
  out->d_prims = b->d_prims;
  out->sort_global_passes = uint32_t(b->cfg.strategy == 2
                                       ? (b->d_seg && seg_pass_applies(b->key_bits, b->cfg)
                                            ? 1
                                            : b->cfg.global_digits)
                                       : b->passes);
  return LBVH_OK;
}

lbvh_status
lbvh_batch_download_scene(lbvh_ctx* c, const lbvh_batch* b, uint32_t scene, void* nodes,
                          uint64_t cap_bytes, uint32_t* codes, uint32_t* ids, float* bounds6)
{
  if (!c || !b)
    return fail(c, LBVH_E_INVALID, "null argument");
  if (scene >= b->nseg)
    return fail(c, LBVH_E_INVALID, "scene index out of range");
  if (ids && !b->with_ids)
    return fail(c, LBVH_E_INVALID, "batch was built without primitive ids (keep_codes)");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  const uint64_t P = b->h_off[scene];
  const uint64_t ns = b->h_off[scene + 1] - P;
  if (nodes) {
    if (cap_bytes < (ns - 1) * 32)
      return fail(c, LBVH_E_INVALID, "destination too small for the node array");
    HIP_TRY(c, hipMemcpyAsync(nodes, static_cast<const char*>(b->d_nodes) + (P - scene) * 32,
                              (ns - 1) * 32, hipMemcpyDeviceToHost, c->stream));
  }
  std::vector<uint64_t> keys;
  if (codes) {
    keys.resize(ns);
    const uint64_t* src = b->sorted_in_b ? b->d_keys_b : b->d_keys_a;
    HIP_TRY(c, hipMemcpyAsync(keys.data(), src + P, ns * 8, hipMemcpyDeviceToHost, c->stream));
  }
  if (ids) {
    const uint32_t* src = b->sorted_in_b ? b->d_ids_b : b->d_ids_a;
    HIP_TRY(c, hipMemcpyAsync(ids, src + P, ns * 4, hipMemcpyDeviceToHost, c->stream));
  }
  int32_t bk[6];
  if (bounds6)
    HIP_TRY(c, hipMemcpyAsync(bk, b->d_bounds_key + 6ull * scene, 24, hipMemcpyDeviceToHost,
                              c->stream));
  st = sync_and_check(c);
  if (st != LBVH_OK)
    return st;
  for (uint64_t k = 0; codes && k < ns; k++)
    codes[k] = uint32_t(keys[k] & kCodeMask);
  for (uint64_t k = 0; ids && k < ns; k++)
    ids[k] = uint32_t(ids[k] - P);
  for (int k = 0; bounds6 && k < 6; k++) {
    const int32_t v = bk[k] >= 0 ? bk[k] : (bk[k] ^ 0x7fffffff);
    std::memcpy(&bounds6[k], &v, 4);
  }
  return LBVH_OK;
}

void
lbvh_batch_destroy(lbvh_batch* b)
{
  free_batch(b);
}

lbvh_status
lbvh_batch_trace_closest_device(lbvh_ctx* c, const lbvh_batch* b, const float* d_rays,
                                const uint64_t* d_ray_off, uint64_t nrays, int mode,
                                uint32_t* d_prim, float* d_t, float* d_uv)
{
  if (!c || !b || (nrays && (!d_rays || !d_ray_off || !d_prim || !d_t)))
    return fail(c, LBVH_E_INVALID, "null argument");
  if (mode < 0 || mode > 2)
    return fail(c, LBVH_E_INVALID, "unknown intersector mode");
  if (nrays > kMaxRaysPerCall)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 rays per call");
  if (misaligned(d_rays, 16) || misaligned(d_ray_off, 8) ||
      trace_outputs_misaligned(d_prim, d_t, d_uv))
    return fail(c, LBVH_E_INVALID, "device rays must be 16-byte aligned, ray offsets and uv "
                                   "8-byte, prim and t 4-byte");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  begin_timed(c, 2);
  record(c, 0);
  HIP_TRY(c, launch_trace(mode, b->d_nodes, b->d_prims, b->n, b->d_off, d_ray_off, b->nseg,
                          d_rays, nrays, d_prim, d_t, d_uv, c->stream));
  record(c, 1, true);
  return LBVH_OK;
}

lbvh_status
lbvh_batch_trace_closest(lbvh_ctx* c, const lbvh_batch* b, const float* rays,
                         const uint64_t* ray_off, int mode, uint32_t* prim, float* t, float* uv)
{
  if (!c || !b || !ray_off)
    return fail(c, LBVH_E_INVALID, "null argument");
  if (ray_off[0] != 0)
    return fail(c, LBVH_E_INVALID, "ray offsets start at 0");
  for (uint32_t s = 0; s < b->nseg; s++)
    if (ray_off[s + 1] < ray_off[s])
      return fail(c, LBVH_E_INVALID, "ray offsets must not decrease");
  const uint64_t nrays = ray_off[b->nseg];
  if (nrays == 0)
    return LBVH_OK;
  if (!rays || !prim || !t)
    return fail(c, LBVH_E_INVALID, "null argument");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  const uint64_t off_bytes = 8ull * (b->nseg + 1);
  const uint64_t off_at = (nrays * 32 + 255) & ~uint64_t(255);
  if ((st = ensure(c, c->io_in, off_at + off_bytes)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->io_out, nrays * 16)) != LBVH_OK)
    return st;
  float* d_rays = static_cast<float*>(c->io_in.p);
  uint64_t* d_off = reinterpret_cast<uint64_t*>(static_cast<char*>(c->io_in.p) + off_at);
  uint32_t* d_prim = static_cast<uint32_t*>(c->io_out.p);
  float* d_t = reinterpret_cast<float*>(d_prim + nrays);
  float* d_uv = d_t + nrays;
  HIP_TRY(c, hipMemcpyAsync(d_rays, rays, nrays * 32, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(d_off, ray_off, off_bytes, hipMemcpyHostToDevice, c->stream));
  st = lbvh_batch_trace_closest_device(c, b, d_rays, d_off, nrays, mode, d_prim, d_t,
                                       uv ? d_uv : nullptr);
  if (st != LBVH_OK)
    return st;
  HIP_TRY(c, hipMemcpyAsync(prim, d_prim, nrays * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(t, d_t, nrays * 4, hipMemcpyDeviceToHost, c->stream));
  if (uv)
    HIP_TRY(c, hipMemcpyAsync(uv, d_uv, nrays * 8, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

} // extern "C"
