// C ABI of lbvh::builder<double> (include/lbvh_mi355x.h): FP64 boxes in,
// node<double> records out, on the context's stream and workspace.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"
#include "lbvh_mi355x.h"
#include "runtime.hpp"

using namespace lbvh_eng;
using namespace lbvh_rt;

namespace {

constexpr uint64_t kMaxF64Prims = (uint64_t(1) << 31) - 1;

lbvh_status queue_f64_build(lbvh_ctx* c, const double* d_boxes, uint64_t n,
                            const lbvh_build_opts& opts, void* d_nodes, const SortConfig& cfg);

// The sort configuration of an FP64 build: explicit options, else the
// context's tuned choice for FP64 builds of this size (a bucket of their
// own, apart from 63-bit FP32 builds of the same key width), tuning now on
// this very build when there is none yet.
lbvh_status
choose_sort_f64(lbvh_ctx* c, const double* d_boxes, uint64_t n, const lbvh_build_opts& opts,
                void* d_nodes, SortConfig* cfg)
{
  const int key_bits = 60;
  if (sort_overridden(opts)) {
    *cfg = sort_from_opts(opts, key_bits, n);
    return LBVH_OK;
  }
  *cfg = default_sort();
  if (!c->autotune || n < kTuneMinPrims)
    return LBVH_OK;
  const uint64_t key = tune_key(false, key_bits, n, 64);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    *cfg = it->second;
    return LBVH_OK;
  }
  float best = -1.0f;
  size_t first = c->tune_log.size();
  for (const SortConfig& cand : sort_candidates(false, key_bits, n)) {
    float ms = -1.0f;
    for (int rep = 0; rep < 2; rep++) {
      lbvh_status st;
      ms = time_call(c, [&] { return queue_f64_build(c, d_boxes, n, opts, d_nodes, cand); }, &st);
      if (st != LBVH_OK)
        return st;
      if ((st = sync_and_check(c)) != LBVH_OK)
        return st;
    }
    c->tune_log.push_back(TuneRecord{ 0, key_bits, ceil_log2_u64(n), cand, ms, false, 64 });
    if (ms >= 0.0f && (best < 0.0f || ms < best)) {
      best = ms;
      *cfg = cand;
    }
  }
  for (size_t i = first; i < c->tune_log.size(); i++)
    c->tune_log[i].picked = same_sort(c->tune_log[i].cfg, *cfg);
  c->tuned[key] = *cfg;
  return LBVH_OK;
}

// Queues the whole FP64 build of n boxes at d_boxes into d_nodes with the
// given sort configuration.
lbvh_status
queue_f64_build(lbvh_ctx* c, const double* d_boxes, uint64_t n, const lbvh_build_opts& opts,
                void* d_nodes, const SortConfig& cfg)
{
  const int key_bits = 60;
  if (!valid_sort_config(cfg, key_bits))
    return fail(c, LBVH_E_INVALID, "unsupported sort digit width / tile");
  const int passes = sort_passes(key_bits, cfg.digit_bits);
  const int first_pass = cfg.strategy == 2 ? passes - cfg.global_digits : 0;
  const uint32_t radix = 1u << cfg.digit_bits;
  const bool with_ids = opts.leaf_order == LBVH_LEAF_PRIMITIVE_ID;

  lbvh_status st;
  if ((st = ensure(c, c->keys_a, 8 * n)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->keys_b, 8 * n)) != LBVH_OK)
    return st;
  if (with_ids) {
    if ((st = ensure(c, c->ids_a, 4 * n)) != LBVH_OK)
      return st;
    if ((st = ensure(c, c->ids_b, 4 * n)) != LBVH_OK)
      return st;
  }
  // bounds partials, then the final bounds
  if ((st = ensure(c, c->partials, 48ull * (f64_bounds_blocks(n) + 1))) != LBVH_OK)
    return st;
  const uint32_t small_words = kSmallHist + uint32_t(passes) * radix;
  if ((st = ensure(c, c->small, 4ull * small_words, true)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->tree, f64_tree_scratch_bytes(n))) != LBVH_OK)
    return st;
  const uint64_t counts_words = sort_counts_words(n, cfg);
  if ((st = ensure(c, c->counts, 4 * counts_words)) != LBVH_OK)
    return st;
  const uint64_t sw = sort_status_words(n, cfg);
  if (c->status.cap < sw * 8) {
    if ((st = ensure(c, c->status, sw * 8, true)) != LBVH_OK)
      return st;
  }
  if (c->epoch_next + kMaxSortPasses >= (1u << 24)) {
    HIP_TRY(c, launch_fill_u32(static_cast<uint32_t*>(c->status.p), c->status.cap / 4, 0u,
                               c->stream));
    c->epoch_next = 1;
  }
  const uint32_t epoch = c->epoch_next;
  c->epoch_next += kMaxSortPasses;

  uint32_t* small = static_cast<uint32_t*>(c->small.p);
  SortWorkspace ws;
  ws.status = static_cast<uint64_t*>(c->status.p);
  ws.status_words = sw;
  ws.tickets = small + kSmallTickets;
  ws.hist = small + kSmallHist;
  ws.err = small + kSmallErr;
  ws.counts = static_cast<uint32_t*>(c->counts.p);
  ws.counts_words = c->counts.cap / 4;
  ws.first_counts = false;

  double* partials = static_cast<double*>(c->partials.p);
  double* bounds = partials + 6 * size_t(f64_bounds_blocks(n));

  begin_timed(c, 1);
  record(c, 0);
  hipError_t e = launch_fill_u32(small + kSmallTickets, small_words - kSmallTickets, 0u, c->stream);
  if (e == hipSuccess)
    e = launch_f64_bounds(d_boxes, n, partials, bounds, c->stream);
  if (e == hipSuccess)
    e = launch_f64_encode(d_boxes, n, bounds, static_cast<uint64_t*>(c->keys_a.p), ws.hist,
                          cfg.digit_bits, first_pass, passes, c->stream);
  record(c, 1);
  bool in_b = false;
  if (e == hipSuccess)
    e = launch_sort(key_bits, 8, n, c->keys_a.p, c->keys_b.p,
                    with_ids ? static_cast<uint32_t*>(c->ids_a.p) : nullptr,
                    with_ids ? static_cast<uint32_t*>(c->ids_b.p) : nullptr, cfg, ws, epoch,
                    &in_b, c->stream);
  record(c, 2);
  const uint64_t* sorted = static_cast<const uint64_t*>(in_b ? c->keys_b.p : c->keys_a.p);
  const uint32_t* ids =
    with_ids ? static_cast<const uint32_t*>(in_b ? c->ids_b.p : c->ids_a.p) : nullptr;
  if (e == hipSuccess)
    e = launch_f64_tree(sorted, ids, d_boxes, n, d_nodes, c->tree.p, c->stream);
  record(c, 3);
  record(c, 4, true);
  if (e != hipSuccess)
    return hip_fail(c, e, "FP64 build launch");
  return LBVH_OK;
}

// Validates the arguments, picks the sort and queues the build.
lbvh_status
f64_build(lbvh_ctx* c, const double* d_boxes, uint64_t n, const lbvh_build_opts* opts_in,
          void* d_nodes)
{
  lbvh_build_opts opts;
  lbvh_build_opts_init(&opts);
  if (opts_in)
    opts = *opts_in;
  if (opts.leaf_order > 1)
    return fail(c, LBVH_E_INVALID, "unknown leaf order");
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  if (n > kMaxF64Prims)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 primitives per FP64 tree");
  if ((reinterpret_cast<uintptr_t>(d_boxes) & 15) != 0 ||
      (reinterpret_cast<uintptr_t>(d_nodes) & 15) != 0)
    return fail(c, LBVH_E_INVALID, "boxes and nodes must be 16-byte aligned");
  SortConfig cfg;
  lbvh_status st = choose_sort_f64(c, d_boxes, n, opts, d_nodes, &cfg);
  if (st != LBVH_OK)
    return st;
  return queue_f64_build(c, d_boxes, n, opts, d_nodes, cfg);
}

// Validates a traversal of a node<double> array and queues it (phase kind 2
// of the context's timing, as FP32 traversals).
lbvh_status
f64_trace(lbvh_ctx* c, const void* d_nodes, uint64_t n, const double* d_tris,
          const double* d_rays, uint64_t nrays, int mode, uint64_t* d_prim, double* d_t,
          double* d_uv)
{
  if (mode < 0 || mode > 1)
    return fail(c, LBVH_E_INVALID, "unknown intersector mode (FP64 traversal: 0 or 1)");
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  if (n > kMaxF64Prims)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 primitives per FP64 tree");
  if (nrays > kMaxRaysPerCall)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 rays per call");
  if (misaligned(d_nodes, 16) || misaligned(d_tris, 16) || misaligned(d_rays, 16) ||
      misaligned(d_prim, 8) || misaligned(d_t, 8) || misaligned(d_uv, 16))
    return fail(c, LBVH_E_INVALID,
                "FP64 traversal: nodes, triangles, rays and uv 16-byte aligned, prim and t 8-byte");
  if (nrays == 0)
    return LBVH_OK;
  begin_timed(c, 2);
  record(c, 0);
  HIP_TRY(c, launch_trace_f64(mode, d_nodes, d_tris, n, d_rays, nrays, d_prim, d_t, d_uv,
                              c->stream));
  record(c, 1, true);
  return LBVH_OK;
}

} // namespace

extern "C" {

lbvh_status
lbvh_trace_closest_f64_device(lbvh_ctx* c, const void* d_nodes, uint64_t n,
                              const double* d_tri96, const double* d_rays64, uint64_t nrays,
                              int mode, uint64_t* d_prim, double* d_t, double* d_uv)
{
  if (!c || !d_nodes || !d_tri96 || (nrays && (!d_rays64 || !d_prim || !d_t)))
    return fail(c, LBVH_E_INVALID, "null argument");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  return f64_trace(c, d_nodes, n, d_tri96, d_rays64, nrays, mode, d_prim, d_t, d_uv);
}

lbvh_status
lbvh_trace_closest_f64(lbvh_ctx* c, const void* nodes, uint64_t n, const double* tri96,
                       const double* rays64, uint64_t nrays, int mode, uint64_t* prim,
                       double* t, double* uv)
{
  if (!c || !nodes || !tri96 || (nrays && (!rays64 || !prim || !t)))
    return fail(c, LBVH_E_INVALID, "null argument");
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  if (n > kMaxF64Prims)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 primitives per FP64 tree");
  if (mode < 0 || mode > 1)
    return fail(c, LBVH_E_INVALID, "unknown intersector mode (FP64 traversal: 0 or 1)");
  if (nrays == 0)
    return LBVH_OK;
  if (nrays > kMaxRaysPerCall)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 rays per call");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  // nodes, triangles and rays share the input buffer, each part 256-aligned
  const uint64_t node_bytes = 64 * (n - 1), tri_bytes = 96 * n, ray_bytes = 64 * nrays;
  const uint64_t tri_off = (node_bytes + 255) & ~uint64_t(255);
  const uint64_t ray_off = tri_off + ((tri_bytes + 255) & ~uint64_t(255));
  if ((st = ensure(c, c->io_in, ray_off + ray_bytes)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->io_out, 32 * nrays)) != LBVH_OK)
    return st;
  char* in = static_cast<char*>(c->io_in.p);
  void* d_nodes = in;
  double* d_tri = reinterpret_cast<double*>(in + tri_off);
  double* d_rays = reinterpret_cast<double*>(in + ray_off);
  uint64_t* d_prim = static_cast<uint64_t*>(c->io_out.p);
  double* d_t = reinterpret_cast<double*>(d_prim + nrays);
  double* d_uv = d_t + nrays;
  HIP_TRY(c, hipMemcpyAsync(d_nodes, nodes, node_bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(d_tri, tri96, tri_bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(d_rays, rays64, ray_bytes, hipMemcpyHostToDevice, c->stream));
  if ((st = f64_trace(c, d_nodes, n, d_tri, d_rays, nrays, mode, d_prim, d_t,
                      uv ? d_uv : nullptr)) != LBVH_OK) {
    (void)hipStreamSynchronize(c->stream);
    return st;
  }
  HIP_TRY(c, hipMemcpyAsync(prim, d_prim, 8 * nrays, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(t, d_t, 8 * nrays, hipMemcpyDeviceToHost, c->stream));
  if (uv)
    HIP_TRY(c, hipMemcpyAsync(uv, d_uv, 16 * nrays, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

lbvh_status
lbvh_build_aabbs_f64_device(lbvh_ctx* c, const double* d_boxes, uint64_t n,
                            const lbvh_build_opts* opts, void* d_nodes)
{
  if (!c || !d_boxes || !d_nodes)
    return fail(c, LBVH_E_INVALID, "null argument");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  return f64_build(c, d_boxes, n, opts, d_nodes);
}

lbvh_status
lbvh_build_aabbs_f64(lbvh_ctx* c, const double* boxes, uint64_t n, const lbvh_build_opts* opts,
                     void* nodes_out)
{
  if (!c || !boxes || !nodes_out)
    return fail(c, LBVH_E_INVALID, "null argument");
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  if (n > kMaxF64Prims)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 primitives per FP64 tree");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  const uint64_t in_bytes = 48 * n, out_bytes = 64 * (n - 1);
  if ((st = ensure(c, c->io_in, in_bytes)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->io_out, out_bytes)) != LBVH_OK)
    return st;
  HIP_TRY(c, hipMemcpyAsync(c->io_in.p, boxes, in_bytes, hipMemcpyHostToDevice, c->stream));
  if ((st = f64_build(c, static_cast<const double*>(c->io_in.p), n, opts, c->io_out.p)) !=
      LBVH_OK) {
    (void)hipStreamSynchronize(c->stream);
    return st;
  }
  HIP_TRY(c, hipMemcpyAsync(nodes_out, c->io_out.p, out_bytes, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

} // extern "C"
