// Allocation-free build (include/lbvh_mi355x.h, lbvh_build_triangles_noalloc).
//
// The same kernels as lbvh_build_triangles_device, with every scratch buffer
// (centroids, key and id ping-pong buffers, sort counts, tree segments and
// tickets, and the parent links the node array is built beside) carved out
// of one file-scope device array of fixed size instead of the context's
// grow-only workspace and a tree handle.  The caller passes the triangles and
// the node array; nothing is allocated per call.  Each device holds its own
// copy of the array for the life of the process; calls that overlap on one
// device share it and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.hpp"
#include "lbvh_mi355x.h"
#include "runtime.hpp"

using namespace lbvh_eng;
using namespace lbvh_rt;

namespace {

// Sized for LBVH_NOALLOC_MAX_PRIMS primitives with either code width and any
// sort configuration this path accepts (the largest layout, 63-bit codes with
// 10-bit digits and 2,048-key tiles, is 18.4 MiB).
constexpr uint64_t kScratchBytes = uint64_t(24) << 20;
constexpr int kMaxDevices = 64;

// A host-side name for the device array: its device address is obtained once
// per device with hipGetSymbolAddress and only that address is used.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

void*
scratch_of(int device)
{
  std::lock_guard<std::mutex> g(g_mu);
  return g_scratch[device];
}

// Byte offsets of one build's buffers inside the scratch array.
struct Layout
{
  uint64_t tickets = 0; // tree tickets, then the small words (contiguous: cleared together)
  uint64_t small = 0;
  uint32_t clear_words = 0; // tree tickets + the error flag (and a spare word)
  uint32_t small_words = 0;
  uint64_t cent = 0, keys_a = 0, keys_b = 0, ids_a = 0, ids_b = 0, partials = 0, tree = 0,
           counts = 0, counts_words = 0, parent = 0, flags = 0, other_end = 0,
           leafprim = 0, bounds = 0, total = 0;
};

uint64_t
align256(uint64_t x)
{
  return (x + 255) & ~uint64_t(255);
}

Layout
layout_for(uint64_t n, int code_bits, const SortConfig& cfg)
{
  const int key_bits = code_bits == 30 ? 30 : 60;
  const uint64_t kw = code_bits == 30 ? 4 : 8;
  const int passes = sort_passes(key_bits, cfg.digit_bits);
  Layout L;
  const uint64_t ticket_words = (tree_ticket_words(n) + 3) & ~uint64_t(3);
  L.small_words = kSmallHist + uint32_t(passes) * (1u << cfg.digit_bits);
  L.clear_words = uint32_t(ticket_words) + kSmallTickets;
  uint64_t o = 0;
  L.tickets = o;
  L.small = o + 4 * ticket_words;
  o = align256(L.small + 4ull * L.small_words);
  auto take = [&](uint64_t bytes) {
    const uint64_t at = o;
    o = align256(o + bytes);
    return at;
  };
  L.cent = take(12 * n);
  L.keys_a = take(kw * n);
  L.keys_b = take(kw * n);
  L.ids_a = take(4 * n);
  L.ids_b = take(4 * n);
  L.partials = take(24ull * centroid_partials_needed(n));
  L.tree = take(tree_scratch_bytes(n));
  L.counts_words = sort_counts_words(n, cfg);
  L.counts = take(4 * L.counts_words);
  L.parent = take(4 * (2 * n - 1));
  L.flags = take(4 * (n - 1));
  L.other_end = take(4 * (n - 1));
  L.leafprim = take(4 * n);
  L.bounds = take(32);
  L.total = o;
  return L;
}

// The sort of an allocation-free build: explicit options, else the context's
// tuned choice for this key width and size when one exists, else the fixed
// default.  No tuning runs here (a tuning run builds trees of its own).
// Onesweep is not offered: its look-back state belongs to a context.
lbvh_status
sort_for(lbvh_ctx* c, const lbvh_build_opts& o, uint64_t n, SortConfig* cfg)
{
  const int key_bits = o.code_bits == 30 ? 30 : 60;
  if (sort_overridden(o)) {
    *cfg = sort_from_opts(o, key_bits, n);
  } else {
    *cfg = default_sort();
    auto it = c->tuned.find(tune_key(false, key_bits, n));
    if (it != c->tuned.end() && it->second.strategy != 0)
      *cfg = it->second;
  }
  if (cfg->strategy == 0)
    return fail(c, LBVH_E_INVALID, "the allocation-free build does not offer the onesweep sort");
  if (!valid_sort_config(*cfg, key_bits))
    return fail(c, LBVH_E_INVALID, "unsupported sort digit width / tile");
  return LBVH_OK;
}

} // namespace

namespace lbvh_rt {

// Looks up this device's copy of the scratch array (the device must be
// current) and clears it once.  Called when a context is created and, if
// that failed, again by the entry point's first call on the device.
lbvh_status
noalloc_init(lbvh_ctx* c)
{
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(c, LBVH_E_INVALID, "device index beyond the allocation-free build's table");
  std::lock_guard<std::mutex> g(g_mu);
  if (g_scratch[c->device])
    return LBVH_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return fail(c, LBVH_E_HIP, "hipGetSymbolAddress of the build scratch failed");
  }
  e = launch_fill_u32(static_cast<uint32_t*>(p), kScratchBytes / 4, 0u, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess)
    return hip_fail(c, e, "clearing the build scratch");
  g_scratch[c->device] = p;
  return LBVH_OK;
}

} // namespace lbvh_rt

extern "C" {

lbvh_status
lbvh_build_triangles_noalloc(lbvh_ctx* c, const float* d_tri, uint64_t n,
                             const lbvh_build_opts* opts_in, void* d_nodes)
{
  if (!c || !d_tri || !d_nodes)
    return fail(c, LBVH_E_INVALID, "null argument");
  if (misaligned(d_tri, 16) || misaligned(d_nodes, 32))
    return fail(c, LBVH_E_INVALID,
                "device triangles must be 16-byte aligned and the node array 32-byte aligned");
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  if (n > LBVH_NOALLOC_MAX_PRIMS)
    return fail(c, LBVH_E_INVALID,
                "the allocation-free build takes at most LBVH_NOALLOC_MAX_PRIMS (262,144) "
                "triangles; use lbvh_build_triangles_device for more");
  lbvh_build_opts opts;
  lbvh_build_opts_init(&opts);
  if (opts_in)
    opts = *opts_in;
  if (opts.code_bits != 30 && opts.code_bits != 63)
    return fail(c, LBVH_E_INVALID, "code_bits must be 30 or 63");
  if (opts.leaf_order > 1)
    return fail(c, LBVH_E_INVALID, "unknown leaf order");
  if (opts.keep_codes)
    return fail(c, LBVH_E_INVALID, "keep_codes needs a tree handle (lbvh_build_triangles_device)");
  SortConfig cfg;
  lbvh_status st = sort_for(c, opts, n, &cfg);
  if (st != LBVH_OK)
    return st;
  const Layout L = layout_for(n, int(opts.code_bits), cfg);
  if (L.total > kScratchBytes)
    return fail(c, LBVH_E_INVALID, "this sort configuration needs more scratch than the "
                                   "allocation-free build holds");
  if ((st = set_device(c)) != LBVH_OK)
    return st;
  if ((st = noalloc_init(c)) != LBVH_OK) // a table lookup once set up
    return st;
  char* base = static_cast<char*>(scratch_of(c->device));
  auto at = [&](uint64_t off) { return static_cast<void*>(base + off); };

  const int key_bits = opts.code_bits == 30 ? 30 : 60;
  const int passes = sort_passes(key_bits, cfg.digit_bits);
  const size_t kw = opts.code_bits == 30 ? 4 : 8;
  const bool with_ids = opts.leaf_order == LBVH_LEAF_PRIMITIVE_ID;
  const uint32_t rts_g = rts_groups(n, cfg, nullptr);
  uint32_t* small = static_cast<uint32_t*>(at(L.small));
  float* cent = static_cast<float*>(at(L.cent));
  float* bounds = static_cast<float*>(at(L.bounds));
  uint32_t* ids_a = static_cast<uint32_t*>(at(L.ids_a));
  uint32_t* ids_b = static_cast<uint32_t*>(at(L.ids_b));
  SortWorkspace ws;
  ws.status = nullptr; // no look-back: reduce-then-scan and hybrid only
  ws.status_words = 0;
  ws.tickets = small + kSmallTickets;
  ws.hist = small + kSmallHist;
  ws.err = small + kSmallErr;
  ws.counts = static_cast<uint32_t*>(at(L.counts));
  ws.counts_words = L.counts_words;
  ws.first_counts = true;

  // The previous call left its tickets at zero unless it failed part way:
  // clear the tree tickets and the error flag.  The centroid-bounds kernel
  // clears the histograms and sort tickets itself.
  hipError_t e = launch_fill_u32(static_cast<uint32_t*>(at(L.tickets)), L.clear_words, 0u,
                                 c->stream);
  if (e == hipSuccess)
    e = launch_centroid_bounds(d_tri, 0, n, cent, cent + n, cent + 2 * n,
                               static_cast<float*>(at(L.partials)), small + kSmallTickets,
                               L.small_words - kSmallTickets, c->stream);
  if (e == hipSuccess)
    e = launch_encode(int(opts.code_bits), cent, cent + n, cent + 2 * n, n,
                      static_cast<const float*>(at(L.partials)), bounds,
                      at(L.keys_a), ws.hist, cfg.digit_bits,
                      cfg.strategy == 2 ? passes - cfg.global_digits : 0, passes, ws.counts,
                      rts_g, rts_chunk_keys(n, cfg), c->stream);
  bool in_b = false;
  if (e == hipSuccess)
    e = launch_sort(key_bits, int(kw), n, at(L.keys_a), at(L.keys_b),
                    with_ids ? ids_a : nullptr, with_ids ? ids_b : nullptr, cfg, ws, 0u, &in_b,
                    c->stream);
  const void* sorted = at(in_b ? L.keys_b : L.keys_a);
  const uint32_t* leaf_ids = with_ids ? (in_b ? ids_b : ids_a) : nullptr;
  if (e == hipSuccess)
    e = launch_tree(int(opts.code_bits), sorted, leaf_ids, n, nullptr, 1, d_tri, 0, d_nodes,
                    static_cast<uint32_t*>(at(L.parent)),
                    with_ids ? static_cast<uint32_t*>(at(L.leafprim)) : nullptr,
                    static_cast<uint32_t*>(at(L.flags)), static_cast<uint32_t*>(at(L.other_end)),
                    at(L.tree), static_cast<uint32_t*>(at(L.tickets)), c->stream);
  if (e != hipSuccess)
    return hip_fail(c, e, "allocation-free build launch");
  return LBVH_OK;
}

} // extern "C"
