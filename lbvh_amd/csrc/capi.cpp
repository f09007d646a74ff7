// Host runtime and C ABI of the gfx950 LBVH engine (include/lbvh_mi355x.h).
//
// A context owns a HIP stream (or uses one the caller installs), a grow-only
// device workspace reused across builds, and optional per-phase HIP-event
// timing.  A bvh handle owns its node array, parent links, fit flags and
// centroid bounds, and either owns a device copy of its primitives (host
// builds) or borrows the caller's device pointer (*_device builds).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"
#include "lbvh_mi355x.h"
#include "runtime.hpp"

using namespace lbvh_eng;
using namespace lbvh_rt;


namespace lbvh_rt {

lbvh_status
fail(lbvh_ctx* c, lbvh_status s, const std::string& msg)
{
  if (c)
    c->last_error = msg;
  return s;
}

lbvh_status
hip_fail(lbvh_ctx* c, hipError_t e, const char* what)
{
  return fail(c, LBVH_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

lbvh_status
ensure(lbvh_ctx* c, DevBuf& b, uint64_t bytes, bool zero_new)
{
  if (bytes == 0)
    bytes = 256;
  if (b.cap >= bytes)
    return LBVH_OK;
  if (b.p) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
      return hip_fail(c, e, "hipStreamSynchronize");
    (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
  }
  uint64_t cap = (bytes + 4095) & ~uint64_t(4095);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, cap);
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return fail(c, LBVH_E_OOM, "device allocation of " + std::to_string(cap) + " bytes failed");
  }
  b.p = p;
  b.cap = cap;
  if (zero_new) {
    hipError_t z = launch_fill_u32(static_cast<uint32_t*>(p), cap / 4, 0u, c->stream);
    if (z != hipSuccess)
      return hip_fail(c, z, "fill");
  }
  return LBVH_OK;
}

void
dev_free(void* p)
{
  if (p)
    (void)hipFree(p);
}

lbvh_status
set_device(lbvh_ctx* c)
{
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess)
    return hip_fail(c, e, "hipSetDevice");
  return LBVH_OK;
}

// Starts the event set of a timed call (no-op when timing is off).
void
begin_timed(lbvh_ctx* c, int kind)
{
  if (!c->timing)
    return;
  if (c->used == c->sets.size()) {
    EventSet es;
    for (auto& e : es.ev)
      (void)hipEventCreate(&e);
    c->sets.push_back(es);
  }
  EventSet& es = c->sets[c->used++];
  es.kind = kind;
  es.recorded = 0;
}

// Records phase boundary `idx` of the current timed call; `final` marks the
// call's last boundary.  Per-call timing keeps only the first and last.
void
record(lbvh_ctx* c, int idx, bool final)
{
  if (!c->timing || c->used == 0 || idx >= 8)
    return;
  EventSet& es = c->sets[c->used - 1];
  if (c->timing == 2) {
    if (idx != 0 && !final)
      return;
    idx = final ? 1 : 0;
  }
  (void)hipEventRecord(es.ev[idx], c->stream);
  es.recorded = idx + 1;
}

SortConfig
default_sort()
{
  SortConfig c;
  c.digit_bits = 8;
  c.items = 16;
  c.strategy = 1;
  return c;
}

int
hybrid_global_digits(int key_bits, int digit_bits, uint64_t n)
{
  const int passes = sort_passes(key_bits, digit_bits);
  for (int gd = 1; gd < passes - 1; gd++) {
    const int hi_bits = key_bits - (passes - gd) * digit_bits;
    if ((n >> hi_bits) <= kHybridMeanBucket)
      return gd;
  }
  return passes - 1;
}

std::vector<SortConfig>
sort_candidates(bool batch, int key_bits, uint64_t n)
{
  std::vector<SortConfig> v;
  // reduce-then-scan at both digit widths and tiles; onesweep only for
  // single builds (its look-back state is tied to an epoch per build, so a
  // replayed graph could not use it); the hybrid where its global passes
  // leave at least two digits to the LDS sort
  for (int strategy = 1; strategy >= (batch ? 1 : 0); strategy--)
    for (int digit : { 8, 10 })
      for (int items : { 8, 16 }) {
        SortConfig c;
        c.digit_bits = digit;
        c.items = items;
        c.strategy = strategy;
        v.push_back(c);
      }
  for (int digit : { 8, 10 })
    for (int items : { 8, 16 }) {
      SortConfig c;
      c.digit_bits = digit;
      c.items = items;
      c.strategy = 2;
      c.global_digits = hybrid_global_digits(key_bits, digit, n);
      if (c.global_digits <= sort_passes(key_bits, digit) - 2)
        v.push_back(c);
    }
  return v;
}

bool
same_sort(const SortConfig& a, const SortConfig& b)
{
  return a.digit_bits == b.digit_bits && a.items == b.items && a.strategy == b.strategy &&
         a.global_digits == b.global_digits;
}

bool
sort_overridden(const lbvh_build_opts& o)
{
  return o.sort_digit_bits != 0 || o.sort_tile != 0 || o.sort_strategy != 0;
}

SortConfig
sort_from_opts(const lbvh_build_opts& o, int key_bits, uint64_t n)
{
  SortConfig c = default_sort();
  if (o.sort_digit_bits)
    c.digit_bits = int(o.sort_digit_bits);
  if (o.sort_tile)
    c.items = int(o.sort_tile / 256u);
  if (o.sort_strategy)
    c.strategy = int(o.sort_strategy) - 1;
  if (c.strategy == 2 && (c.digit_bits == 8 || c.digit_bits == 10))
    c.global_digits = hybrid_global_digits(key_bits, c.digit_bits, n);
  return c;
}

int
ceil_log2_u64(uint64_t x)
{
  int b = 0;
  while ((uint64_t(1) << b) < x)
    b++;
  return b;
}

uint64_t
tune_key(bool batch, int key_bits, uint64_t n, int scalar_bits)
{
  return (uint64_t(scalar_bits == 64) << 48) | (uint64_t(batch) << 40) |
         (uint64_t(key_bits) << 32) | uint64_t(ceil_log2_u64(n));
}

bool
valid_sort_config(const SortConfig& c, int key_bits)
{
  if (c.strategy < 0 || c.strategy > 2)
    return false;
  if (c.strategy == 2 && (c.global_digits < 1 ||
                          c.global_digits >= sort_passes(key_bits, c.digit_bits)))
    return false;
  if (c.strategy >= 1)
    return (c.digit_bits == 8 || c.digit_bits == 10) && (c.items == 8 || c.items == 16);
  if (c.digit_bits == 8)
    return c.items == 4 || c.items == 8 || c.items == 16;
  if (c.digit_bits == 10)
    return c.items == 8 || c.items == 16;
  return false;
}

lbvh_status
alloc_bvh(lbvh_ctx* c, lbvh_bvh* b, bool keep_codes)
{
  const uint64_t n = b->n;
  const uint64_t m = n - 1;
  size_t cw = b->code_bits == 30 ? 4 : 8;
  struct
  {
    void** p;
    uint64_t bytes;
  } req[] = {
    { &b->d_nodes, m * 32 },
    { reinterpret_cast<void**>(&b->d_parent), (2 * n - 1) * 4 },
    { reinterpret_cast<void**>(&b->d_flags), m * 4 },
    { reinterpret_cast<void**>(&b->d_other_end), m * 4 },
    { reinterpret_cast<void**>(&b->d_bounds), 32 },
    { &b->d_codes, keep_codes ? n * cw : 0 },
    { reinterpret_cast<void**>(&b->d_ids), keep_codes ? n * 4 : 0 },
    { reinterpret_cast<void**>(&b->d_leafprim), b->leaf_order == LBVH_LEAF_PRIMITIVE_ID ? n * 4 : 0 },
  };
  // one allocation for all of them (256-byte aligned parts): an allocation
  // costs host time on the submitting thread, whatever its size
  uint64_t total = 0;
  for (auto& r : req)
    if (r.bytes)
      total += (r.bytes + 255) & ~uint64_t(255);
  hipError_t e = hipMalloc(&b->block, total);
  if (e != hipSuccess || !b->block) {
    (void)hipGetLastError();
    b->block = nullptr;
    return fail(c, LBVH_E_OOM, "device allocation for the tree failed");
  }
  char* at = static_cast<char*>(b->block);
  for (auto& r : req) {
    if (!r.bytes)
      continue;
    *r.p = at;
    at += (r.bytes + 255) & ~uint64_t(255);
  }
  b->bytes += total;
  return LBVH_OK;
}

void
free_bvh(lbvh_bvh* b)
{
  if (!b)
    return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(b->device);
  dev_free(b->block); // every tree buffer
  if (b->owns_prims)
    dev_free(b->d_prims);
  (void)hipSetDevice(cur);
  delete b;
}

lbvh_status build_common(lbvh_ctx* c, const float* d_prims, int kind, uint64_t n,
                         const lbvh_build_opts* opts_in, bool owns, lbvh_bvh** out,
                         const SortConfig* forced = nullptr);

lbvh_status
reserve_build(lbvh_ctx* c, uint64_t n, int code_bits, const SortConfig& cfg, bool with_ids)
{
  const int key_bits = code_bits == 30 ? 30 : 60;
  const int passes = sort_passes(key_bits, cfg.digit_bits);
  const uint64_t kw = code_bits == 30 ? 4 : 8;
  lbvh_status st;
  if ((st = ensure(c, c->cent, 12 * n)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->keys_a, kw * n)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->keys_b, kw * n)) != LBVH_OK)
    return st;
  if (with_ids) {
    if ((st = ensure(c, c->ids_a, 4 * n)) != LBVH_OK)
      return st;
    if ((st = ensure(c, c->ids_b, 4 * n)) != LBVH_OK)
      return st;
  }
  if ((st = ensure(c, c->partials, 24ull * centroid_partials_needed(n))) != LBVH_OK)
    return st;
  const uint32_t small_words = kSmallHist + uint32_t(passes) * (1u << cfg.digit_bits);
  if ((st = ensure(c, c->small, 4ull * small_words, true)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->tree, tree_scratch_bytes(n))) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->tree_tickets, 4 * tree_ticket_words(n), true)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->counts, 4 * sort_counts_words(n, cfg))) != LBVH_OK)
    return st;
  const uint64_t sw = sort_status_words(n, cfg);
  if (c->status.cap < sw * 8) {
    if ((st = ensure(c, c->status, sw * 8, true)) != LBVH_OK)
      return st;
  }
  return LBVH_OK;
}

SortConfig
untuned_sort(int key_bits, uint64_t n)
{
  SortConfig c = default_sort();
  if (n < kTuneMinPrims)
    return c;
  // the hybrid at 10-bit digits and 4,096-key tiles: the tuner's pick at 1M
  // and 16M 63-bit keys and for batches
  SortConfig h;
  h.digit_bits = 10;
  h.items = 16;
  h.strategy = 2;
  h.global_digits = hybrid_global_digits(key_bits, h.digit_bits, n);
  return h.global_digits <= sort_passes(key_bits, h.digit_bits) - 2 ? h : c;
}

// The sort configuration of a single-scene build: explicit options, else
// the context's tuned choice for this key width and size, tuning now (every
// candidate timed on this very build, twice) when there is none yet.
lbvh_status
choose_sort(lbvh_ctx* c, const float* d_prims, int kind, uint64_t n, const lbvh_build_opts& opts,
            SortConfig* cfg)
{
  const int key_bits = opts.code_bits == 30 ? 30 : 60;
  if (sort_overridden(opts)) {
    *cfg = sort_from_opts(opts, key_bits, n);
    return LBVH_OK;
  }
  *cfg = untuned_sort(key_bits, n);
  if (!c->autotune || n < kTuneMinPrims)
    return LBVH_OK;
  const uint64_t key = tune_key(false, key_bits, n);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    *cfg = it->second;
    return LBVH_OK;
  }
  float best = -1.0f;
  size_t first = c->tune_log.size();
  for (const SortConfig& cand : sort_candidates(false, key_bits, n)) {
    float ms = -1.0f;
    for (int rep = 0; rep < 2; rep++) { // the faster of two runs
      lbvh_bvh* tmp = nullptr;
      lbvh_status st;
      const float r = time_call(c, [&] { return build_common(c, d_prims, kind, n, &opts, false,
                                                             &tmp, &cand); }, &st);
      if (tmp)
        free_bvh(tmp);
      if (st != LBVH_OK)
        return st;
      if ((st = sync_and_check(c)) != LBVH_OK)
        return st;
      if (r >= 0.0f && (ms < 0.0f || r < ms))
        ms = r;
    }
    c->tune_log.push_back(TuneRecord{ 0, key_bits, ceil_log2_u64(n), cand, ms, false });
    if (ms >= 0.0f && (best < 0.0f || ms < best)) {
      best = ms;
      *cfg = cand;
    }
  }
  for (size_t i = first; i < c->tune_log.size(); i++) {
    c->tune_log[i].picked = same_sort(c->tune_log[i].cfg, *cfg);
  }
  c->tuned[key] = *cfg;
  return LBVH_OK;
}

lbvh_status
build_common(lbvh_ctx* c, const float* d_prims, int kind, uint64_t n,
             const lbvh_build_opts* opts_in, bool owns, lbvh_bvh** out, const SortConfig* forced)
{
  lbvh_build_opts opts;
  lbvh_build_opts_init(&opts);
  if (opts_in)
    opts = *opts_in;
  if (opts.code_bits != 30 && opts.code_bits != 63)
    return fail(c, LBVH_E_INVALID, "code_bits must be 30 or 63");
  if (opts.leaf_order > 1)
    return fail(c, LBVH_E_INVALID, "unknown leaf order");
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  if (n >= (uint64_t(1) << 29))
    return fail(c, LBVH_E_INVALID, "at most 2^29 - 1 primitives per tree");
  SortConfig cfg;
  if (forced) {
    cfg = *forced;
  } else {
    lbvh_status st0 = choose_sort(c, d_prims, kind, n, opts, &cfg);
    if (st0 != LBVH_OK)
      return st0;
  }
  const int key_bits = opts.code_bits == 30 ? 30 : 60;
  if (!valid_sort_config(cfg, key_bits))
    return fail(c, LBVH_E_INVALID, "unsupported sort digit width / tile");
  const int passes = sort_passes(key_bits, cfg.digit_bits);
  const uint32_t radix = 1u << cfg.digit_bits;
  const size_t kw = opts.code_bits == 30 ? 4 : 8;
  const bool with_ids = opts.leaf_order == LBVH_LEAF_PRIMITIVE_ID || opts.keep_codes;

  lbvh_status st = reserve_build(c, n, int(opts.code_bits), cfg, with_ids);
  if (st != LBVH_OK)
    return st;
  // reduce-then-scan and hybrid sorts: the encode counts the first pass's
  // chunks itself (one encode workgroup per chunk)
  const bool chunked = cfg.strategy == 1 || cfg.strategy == 2;
  const uint32_t rts_g = chunked ? rts_groups(n, cfg, nullptr) : 0u;
  const uint32_t small_words = kSmallHist + uint32_t(passes) * radix;
  const uint64_t sw = sort_status_words(n, cfg);
  if (c->epoch_next + kMaxSortPasses >= (1u << 24)) {
    HIP_TRY(c, launch_fill_u32(static_cast<uint32_t*>(c->status.p), c->status.cap / 4, 0u,
                               c->stream));
    c->epoch_next = 1;
  }
  const uint32_t epoch = c->epoch_next;
  c->epoch_next += kMaxSortPasses;

  auto* b = new (std::nothrow) lbvh_bvh;
  static std::atomic<uint64_t> next_tree_id{ 1 };
  if (b)
    b->id = next_tree_id.fetch_add(1);
  if (!b)
    return fail(c, LBVH_E_OOM, "host allocation failed");
  b->device = c->device;
  b->n = n;
  b->code_bits = opts.code_bits;
  b->leaf_order = opts.leaf_order;
  b->prim_kind = kind;
  b->sort = cfg;
  b->d_prims = const_cast<float*>(d_prims);
  b->owns_prims = owns;
  if ((st = alloc_bvh(c, b, opts.keep_codes != 0)) != LBVH_OK) {
    b->owns_prims = false;
    free_bvh(b);
    return st;
  }
  if (owns)
    b->bytes += n * (kind == 0 ? 48 : 24);

  uint32_t* small = static_cast<uint32_t*>(c->small.p);
  float* cent = static_cast<float*>(c->cent.p);
  SortWorkspace ws;
  ws.status = static_cast<uint64_t*>(c->status.p);
  ws.status_words = sw;
  ws.tickets = small + kSmallTickets;
  ws.hist = small + kSmallHist;
  ws.err = small + kSmallErr;
  ws.counts = static_cast<uint32_t*>(c->counts.p);
  ws.counts_words = c->counts.cap / 4;
  ws.first_counts = chunked;

  begin_timed(c, 1);
  record(c, 0);
  hipError_t e = launch_centroid_bounds(d_prims, kind, n, cent, cent + n, cent + 2 * n,
                                        static_cast<float*>(c->partials.p),
                                        small + kSmallTickets, small_words - kSmallTickets,
                                        c->stream);
  if (e == hipSuccess)
    e = launch_encode(opts.code_bits, cent, cent + n, cent + 2 * n, n,
                      static_cast<const float*>(c->partials.p), b->d_bounds,
                      c->keys_a.p, ws.hist,
                      cfg.digit_bits, cfg.strategy == 2 ? passes - cfg.global_digits : 0, passes,
                      chunked ? ws.counts : nullptr, rts_g,
                      chunked ? rts_chunk_keys(n, cfg) : 0, c->stream);
  record(c, 1);
  bool in_b = false;
  if (e == hipSuccess)
    e = launch_sort(key_bits, int(kw), n, c->keys_a.p, c->keys_b.p,
                    with_ids ? static_cast<uint32_t*>(c->ids_a.p) : nullptr,
                    with_ids ? static_cast<uint32_t*>(c->ids_b.p) : nullptr, cfg, ws, epoch,
                    &in_b, c->stream);
  record(c, 2);
  void* sorted = in_b ? c->keys_b.p : c->keys_a.p;
  uint32_t* sorted_ids = with_ids ? static_cast<uint32_t*>(in_b ? c->ids_b.p : c->ids_a.p)
                                  : nullptr;
  // tree and fit in one kernel (the fit phase is part of the tree phase)
  const uint32_t* leaf_ids = opts.leaf_order == LBVH_LEAF_PRIMITIVE_ID ? sorted_ids : nullptr;
  if (e == hipSuccess)
    e = launch_tree(opts.code_bits, sorted, leaf_ids, n, nullptr, 1, d_prims, kind, b->d_nodes,
                    b->d_parent, b->d_leafprim, b->d_flags, b->d_other_end, c->tree.p,
                    static_cast<uint32_t*>(c->tree_tickets.p), c->stream);
  record(c, 3);
  record(c, 4, true);
  if (e == hipSuccess && opts.keep_codes) {
    e = hipMemcpyAsync(b->d_codes, sorted, kw * n, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->d_ids, sorted_ids, 4 * n, hipMemcpyDeviceToDevice, c->stream);
  }
  if (e != hipSuccess) {
    b->owns_prims = false;
    free_bvh(b);
    return hip_fail(c, e, "build launch");
  }
  *out = b;
  return LBVH_OK;
}

// Reads and clears the bounded-wait error flag; synchronizes the stream.
lbvh_status
sync_and_check(lbvh_ctx* c)
{
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->small.p) {
    uint32_t err = 0;
    HIP_TRY(c, hipMemcpy(&err, static_cast<uint32_t*>(c->small.p) + kSmallErr, 4,
                         hipMemcpyDeviceToHost));
    if (err) {
      HIP_TRY(c, launch_fill_u32(static_cast<uint32_t*>(c->small.p) + kSmallErr, 1, 0u,
                                 c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      return fail(c, LBVH_E_TIMEOUT, "a radix-sort look-back wait exceeded its spin limit");
    }
  }
  return LBVH_OK;
}

} // namespace lbvh_rt

extern "C" {

const char*
lbvh_status_string(lbvh_status s)
{
  switch (s) {
    case LBVH_OK:
      return "ok";
    case LBVH_E_INVALID:
      return "invalid argument";
    case LBVH_E_OOM:
      return "out of device memory";
    case LBVH_E_HIP:
      return "HIP runtime error";
    case LBVH_E_BUSY:
      return "busy";
    case LBVH_E_NODEVICE:
      return "no usable gfx950 device";
    case LBVH_E_TIMEOUT:
      return "bounded wait expired";
  }
  return "unknown status";
}

void
lbvh_build_opts_init(lbvh_build_opts* o)
{
  if (!o)
    return;
  std::memset(o, 0, sizeof(*o));
  o->code_bits = 30;
  o->leaf_order = LBVH_LEAF_REFERENCE;
}

lbvh_status
lbvh_ctx_create(int device, lbvh_ctx** out)
{
  if (!out)
    return LBVH_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return LBVH_E_NODEVICE;
  }
  if (device < 0 || device >= count)
    return LBVH_E_INVALID;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return LBVH_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return LBVH_E_NODEVICE;
  auto* c = new (std::nothrow) lbvh_ctx;
  if (!c)
    return LBVH_E_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream,
                                                                      hipStreamNonBlocking) !=
                                               hipSuccess) {
    delete c;
    return LBVH_E_HIP;
  }
  c->stream = c->own_stream;
  // set-up of the allocation-free build on this device; if it fails here,
  // that entry point retries on its first call and reports the error
  if (noalloc_init(c) != LBVH_OK)
    c->last_error.clear();
  *out = c;
  return LBVH_OK;
}

void
lbvh_ctx_destroy(lbvh_ctx* c)
{
  if (!c)
    return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = { &c->cent, &c->keys_a, &c->keys_b, &c->ids_a, &c->ids_b, &c->partials,
                     &c->small, &c->status, &c->counts, &c->treelets,
                     &c->io_in, &c->io_out, &c->tree, &c->tree_tickets, &c->treelet_words,
                     &c->ray_keys, &c->ray_ids, &c->ray_counts, &c->ray_hist };
  for (auto* b : bufs)
    dev_free(b->p);
  for (auto& es : c->sets)
    for (auto& e : es.ev)
      if (e)
        (void)hipEventDestroy(e);
  if (c->own_stream)
    (void)hipStreamDestroy(c->own_stream);
  delete c;
}

lbvh_status
lbvh_ctx_set_stream(lbvh_ctx* c, void* stream)
{
  if (!c)
    return LBVH_E_INVALID;
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
  return LBVH_OK;
}

lbvh_status
lbvh_ctx_synchronize(lbvh_ctx* c)
{
  if (!c)
    return LBVH_E_INVALID;
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  return sync_and_check(c);
}

lbvh_status
lbvh_ctx_set_timing(lbvh_ctx* c, int enable)
{
  if (!c)
    return LBVH_E_INVALID;
  if (enable < 0 || enable > 2)
    return fail(c, LBVH_E_INVALID, "timing mode is 0, 1 or 2");
  c->timing = enable;
  c->used = 0;
  return LBVH_OK;
}

// Sums the phase times of timed calls of `kind` (0 = the last call only).
static lbvh_status
phase_sum(lbvh_ctx* c, int kind, float* ms, uint32_t cap, uint32_t* count, uint32_t* calls)
{
  if (!c)
    return LBVH_E_INVALID;
  if (count)
    *count = 0;
  if (calls)
    *calls = 0;
  if (c->used == 0)
    return LBVH_OK;
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  size_t first = kind == 0 ? c->used - 1 : 0;
  uint32_t nph = 0, ncalls = 0;
  for (uint32_t i = 0; i < cap; i++)
    ms[i] = 0.0f;
  for (size_t s = first; s < c->used; s++) {
    EventSet& es = c->sets[s];
    if (kind != 0 && es.kind != kind)
      continue;
    uint32_t k = es.recorded > 1 ? uint32_t(es.recorded - 1) : 0u;
    if (k == 0)
      continue;
    HIP_TRY(c, hipEventSynchronize(es.ev[k]));
    for (uint32_t i = 0; i < k && i < cap; i++) {
      float v = 0.0f;
      HIP_TRY(c, hipEventElapsedTime(&v, es.ev[i], es.ev[i + 1]));
      ms[i] += v;
    }
    if (k > nph)
      nph = k;
    ncalls++;
  }
  if (count)
    *count = nph;
  if (calls)
    *calls = ncalls;
  return LBVH_OK;
}

lbvh_status
lbvh_ctx_phase_ms(lbvh_ctx* c, float* ms, uint32_t cap, uint32_t* count)
{
  return phase_sum(c, 0, ms, cap, count, nullptr);
}

lbvh_status
lbvh_ctx_phase_stats(lbvh_ctx* c, int kind, float* sum_ms, uint32_t cap, uint32_t* count,
                     uint32_t* calls)
{
  if (kind < 1 || kind > 3)
    return fail(c, LBVH_E_INVALID, "kind is 1 (build), 2 (trace) or 3 (refit)");
  return phase_sum(c, kind, sum_ms, cap, count, calls);
}

const char*
lbvh_ctx_last_error(const lbvh_ctx* c)
{
  return c ? c->last_error.c_str() : "null context";
}

uint64_t
lbvh_ctx_workspace_bytes(const lbvh_ctx* c)
{
  if (!c)
    return 0;
  return c->cent.cap + c->keys_a.cap + c->keys_b.cap + c->ids_a.cap + c->ids_b.cap +
         c->partials.cap + c->small.cap + c->status.cap + c->counts.cap +
         c->treelets.cap + c->io_in.cap + c->io_out.cap + c->tree.cap + c->tree_tickets.cap +
         c->treelet_words.cap + c->ray_keys.cap + c->ray_ids.cap + c->ray_counts.cap +
         c->ray_hist.cap;
}

lbvh_status
lbvh_build_triangles_device(lbvh_ctx* c, const float* d_tri, uint64_t n,
                            const lbvh_build_opts* opts, lbvh_bvh** out)
{
  if (!c || !out || (!d_tri && n))
    return fail(c, LBVH_E_INVALID, "null argument");
  *out = nullptr;
  if (misaligned(d_tri, 16))
    return fail(c, LBVH_E_INVALID, "device triangles must be 16-byte aligned");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  return build_common(c, d_tri, 0, n, opts, false, out);
}

lbvh_status
lbvh_build_aabbs_device(lbvh_ctx* c, const float* d_boxes, uint64_t n,
                        const lbvh_build_opts* opts, lbvh_bvh** out)
{
  if (!c || !out || (!d_boxes && n))
    return fail(c, LBVH_E_INVALID, "null argument");
  *out = nullptr;
  if (misaligned(d_boxes, 8))
    return fail(c, LBVH_E_INVALID, "device boxes must be 8-byte aligned");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  return build_common(c, d_boxes, 1, n, opts, false, out);
}

static lbvh_status
build_host(lbvh_ctx* c, const float* host, int kind, uint64_t n, const lbvh_build_opts* opts,
           lbvh_bvh** out)
{
  if (!c || !out || (!host && n))
    return fail(c, LBVH_E_INVALID, "null argument");
  *out = nullptr;
  if (n < 2)
    return fail(c, LBVH_E_INVALID, "a tree needs at least 2 primitives");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  const uint64_t bytes = n * (kind == 0 ? 48 : 24);
  float* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, LBVH_E_OOM, "device allocation for primitives failed");
  }
  hipError_t e = hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) {
    dev_free(d);
    return hip_fail(c, e, "hipMemcpyAsync");
  }
  st = build_common(c, d, kind, n, opts, true, out);
  if (st != LBVH_OK) {
    (void)hipStreamSynchronize(c->stream);
    dev_free(d);
    return st;
  }
  st = sync_and_check(c);
  if (st != LBVH_OK) {
    free_bvh(*out);
    *out = nullptr;
  }
  return st;
}

lbvh_status
lbvh_build_triangles(lbvh_ctx* c, const float* tri, uint64_t n, const lbvh_build_opts* opts,
                     lbvh_bvh** out)
{
  return build_host(c, tri, 0, n, opts, out);
}

lbvh_status
lbvh_build_aabbs(lbvh_ctx* c, const float* boxes, uint64_t n, const lbvh_build_opts* opts,
                 lbvh_bvh** out)
{
  return build_host(c, boxes, 1, n, opts, out);
}

// Queues a refit of `b` over its current primitives.  The treelet list
// depends on the topology only, so a context keeps the list of the tree it
// last refitted and a refit of the same tree reuses it.
static lbvh_status
queue_refit(lbvh_ctx* c, lbvh_bvh* b)
{
  lbvh_status st;
  const void* kept = c->treelets.p;
  if ((st = ensure(c, c->treelets, 8 * b->n)) != LBVH_OK)
    return st;
  if (c->treelets.p != kept)
    c->treelets_of = 0; // reallocated: nothing kept
  if ((st = ensure(c, c->treelet_words, 8, true)) != LBVH_OK)
    return st;
  uint32_t* tw = static_cast<uint32_t*>(c->treelet_words.p); // count, pull queue
  const bool enumerate = c->treelets_of != b->id;
  begin_timed(c, 3);
  record(c, 0);
  HIP_TRY(c, launch_fill_u32(enumerate ? tw : tw + 1, enumerate ? 2 : 1, 0u, c->stream));
  if (enumerate) {
    c->treelets_of = 0;
    HIP_TRY(c, launch_enum_treelets(b->n, nullptr, 1, b->d_parent, b->d_other_end, c->treelets.p,
                                    tw, c->stream));
    c->treelets_of = b->id;
  }
  HIP_TRY(c, launch_fit(b->d_prims, 0, b->n, nullptr, 1, b->d_nodes, b->d_parent, b->d_other_end,
                        b->d_leafprim, b->d_flags, c->treelets.p, tw, tw + 1, c->stream));
  record(c, 1, true);
  return LBVH_OK;
}

lbvh_status
lbvh_refit_triangles_device(lbvh_ctx* c, lbvh_bvh* b, const float* d_tri)
{
  if (!c || !b || !d_tri)
    return fail(c, LBVH_E_INVALID, "null argument");
  if (b->prim_kind != 0)
    return fail(c, LBVH_E_INVALID, "refit_triangles on a tree built from boxes");
  if (misaligned(d_tri, 16))
    return fail(c, LBVH_E_INVALID, "device triangles must be 16-byte aligned");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if (b->owns_prims) {
    HIP_TRY(c, hipMemcpyAsync(b->d_prims, d_tri, b->n * 48, hipMemcpyDeviceToDevice, c->stream));
  } else {
    b->d_prims = const_cast<float*>(d_tri);
  }
  return queue_refit(c, b);
}

lbvh_status
lbvh_refit_triangles(lbvh_ctx* c, lbvh_bvh* b, const float* tri)
{
  if (!c || !b || !tri)
    return fail(c, LBVH_E_INVALID, "null argument");
  if (b->prim_kind != 0)
    return fail(c, LBVH_E_INVALID, "refit_triangles on a tree built from boxes");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if (!b->owns_prims) {
    float* d = nullptr;
    if (hipMalloc(&d, b->n * 48) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, LBVH_E_OOM, "device allocation for primitives failed");
    }
    b->d_prims = d;
    b->owns_prims = true;
    b->bytes += b->n * 48;
  }
  HIP_TRY(c, hipMemcpyAsync(b->d_prims, tri, b->n * 48, hipMemcpyHostToDevice, c->stream));
  if ((st = queue_refit(c, b)) != LBVH_OK)
    return st;
  return sync_and_check(c);
}

lbvh_status
lbvh_bvh_get_info(const lbvh_bvh* b, lbvh_bvh_info* out)
{
  if (!b || !out)
    return LBVH_E_INVALID;
  out->n_prims = b->n;
  out->n_nodes = b->n - 1;
  out->code_bits = b->code_bits;
  out->leaf_order = b->leaf_order;
  out->d_nodes = b->d_nodes;
  out->d_prims = b->prim_kind == 0 ? b->d_prims : nullptr;
  out->device_bytes = b->bytes;
  out->sort_digit_bits = uint32_t(b->sort.digit_bits);
  out->sort_tile = uint32_t(b->sort.items * 256);
  out->sort_strategy = uint32_t(b->sort.strategy + 1);
  out->sort_global_digits = uint32_t(b->sort.global_digits);
  return LBVH_OK;
}

lbvh_status
lbvh_ctx_set_autotune(lbvh_ctx* c, int enable)
{
  if (!c)
    return LBVH_E_INVALID;
  c->autotune = enable != 0;
  return LBVH_OK;
}

uint32_t
lbvh_ctx_tune_log(const lbvh_ctx* c, lbvh_tune_record* out, uint32_t cap)
{
  if (!c)
    return 0;
  const uint32_t n = uint32_t(c->tune_log.size());
  for (uint32_t i = 0; out && i < n && i < cap; i++) {
    const TuneRecord& r = c->tune_log[i];
    const bool sort = r.kind == 0;
    out[i].batch = uint32_t(r.batch);
    out[i].key_bits = uint32_t(r.key_bits);
    out[i].log2_n = uint32_t(r.log2n);
    out[i].sort_digit_bits = sort ? uint32_t(r.cfg.digit_bits) : 0u;
    out[i].sort_tile = sort ? uint32_t(r.cfg.items * 256) : 0u;
    out[i].sort_strategy = sort ? uint32_t(r.cfg.strategy + 1) : 0u;
    out[i].picked = r.picked ? 1u : 0u;
    out[i].ms = r.ms;
    out[i].scalar_bits = uint32_t(r.scalar_bits);
    out[i].kind = uint32_t(r.kind);
    out[i].trace_reorder = uint32_t(r.trace.reorder);
    out[i].trace_waves_per_simd = uint32_t(r.trace.waves_per_simd);
    out[i].trace_mode = uint32_t(r.trace_mode);
  }
  return n;
}

lbvh_status
lbvh_bvh_download_nodes(lbvh_ctx* c, const lbvh_bvh* b, void* dst, uint64_t cap)
{
  if (!c || !b || !dst)
    return fail(c, LBVH_E_INVALID, "null argument");
  const uint64_t bytes = (b->n - 1) * 32;
  if (cap < bytes)
    return fail(c, LBVH_E_INVALID, "destination too small for the node array");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  HIP_TRY(c, hipMemcpyAsync(dst, b->d_nodes, bytes, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

lbvh_status
lbvh_bvh_download_codes(lbvh_ctx* c, const lbvh_bvh* b, void* codes, uint32_t* ids,
                        float* bounds6)
{
  if (!c || !b)
    return fail(c, LBVH_E_INVALID, "null argument");
  if ((codes || ids) && !b->d_codes)
    return fail(c, LBVH_E_INVALID, "tree was built without keep_codes");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if (codes)
    HIP_TRY(c, hipMemcpyAsync(codes, b->d_codes, b->n * (b->code_bits == 30 ? 4 : 8),
                              hipMemcpyDeviceToHost, c->stream));
  if (ids)
    HIP_TRY(c, hipMemcpyAsync(ids, b->d_ids, b->n * 4, hipMemcpyDeviceToHost, c->stream));
  if (bounds6)
    HIP_TRY(c, hipMemcpyAsync(bounds6, b->d_bounds, 24, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

void
lbvh_bvh_destroy(lbvh_bvh* b)
{
  free_bvh(b);
}

} // extern "C"

namespace {

// Traversal candidates of the tuner: rays in key order or as given, and
// waves per SIMD (8: as many as the registers allow).
std::vector<TraceConfig>
trace_candidates()
{
  std::vector<TraceConfig> v;
  for (int reorder : { 0, 1 })
    for (int waves : { 8, 6, 4 }) {
      TraceConfig t;
      t.reorder = reorder;
      t.waves_per_simd = waves;
      v.push_back(t);
    }
  return v;
}

uint64_t
trace_tune_key(int mode, uint64_t nrays)
{
  return (uint64_t(1) << 56) | (uint64_t(mode) << 40) | uint64_t(ceil_log2_u64(nrays));
}

// Traversals below this many rays are launch-bound: the default is used.
constexpr uint64_t kTraceTuneMinRays = 1u << 14;

// Queues one closest-hit traversal with configuration tc (timed as one call).
lbvh_status
queue_trace(lbvh_ctx* c, const lbvh_bvh* b, const float* d_rays, uint64_t nrays, int mode,
            uint32_t* d_prim, float* d_t, float* d_uv, const TraceConfig& tc)
{
  const uint32_t* order = nullptr;
  lbvh_status st;
  SortConfig sc;
  sc.digit_bits = 8;
  sc.items = 16;
  sc.strategy = 1; // reduce-then-scan: two passes, no look-back state
  const bool reorder = tc.reorder && nrays >= 2 && nrays < (uint64_t(1) << 31);
  if (reorder) {
    const uint32_t groups = rts_groups(nrays, sc, nullptr);
    const uint64_t chunk = rts_chunk_keys(nrays, sc);
    const uint64_t counts_words = sort_counts_words(nrays, sc);
    if ((st = ensure(c, c->ray_keys, 8 * nrays)) != LBVH_OK ||
        (st = ensure(c, c->ray_ids, 8 * nrays)) != LBVH_OK ||
        (st = ensure(c, c->ray_counts, 4 * counts_words)) != LBVH_OK ||
        (st = ensure(c, c->ray_hist, 4 * 512)) != LBVH_OK)
      return st;
    uint32_t* keys = static_cast<uint32_t*>(c->ray_keys.p);
    uint32_t* ids = static_cast<uint32_t*>(c->ray_ids.p);
    SortWorkspace ws;
    ws.hist = static_cast<uint32_t*>(c->ray_hist.p);
    ws.counts = static_cast<uint32_t*>(c->ray_counts.p);
    ws.counts_words = c->ray_counts.cap / 4;
    ws.first_counts = true;
    begin_timed(c, 2);
    record(c, 0);
    HIP_TRY(c, launch_fill_u32(ws.hist, 512, 0u, c->stream));
    HIP_TRY(c, launch_ray_keys(d_rays, nrays, b->d_bounds, keys, ws.hist, ws.counts, groups, chunk,
                               c->stream));
    bool in_b = false;
    HIP_TRY(c, launch_sort(kRayKeyBits, 4, nrays, keys, keys + nrays, ids, ids + nrays, sc, ws, 0u,
                           &in_b, c->stream));
    order = in_b ? ids + nrays : ids;
  } else {
    begin_timed(c, 2);
    record(c, 0);
  }
  HIP_TRY(c, launch_trace(mode, b->d_nodes, b->d_prims, b->n, nullptr, nullptr, 1, d_rays, nrays,
                          d_prim, d_t, d_uv, c->stream, order, tc.waves_per_simd));
  record(c, 1, true);
  return LBVH_OK;
}

// The traversal configuration for this intersector mode and batch size:
// the context's tuned choice, tuning now (every candidate timed on this very
// call, twice) when there is none yet.  Results are identical for every
// candidate.
lbvh_status
choose_trace(lbvh_ctx* c, const lbvh_bvh* b, const float* d_rays, uint64_t nrays, int mode,
             uint32_t* d_prim, float* d_t, float* d_uv, TraceConfig* tc)
{
  *tc = TraceConfig();
  if (!c->autotune || nrays < kTraceTuneMinRays || nrays >= (uint64_t(1) << 31))
    return LBVH_OK;
  const uint64_t key = trace_tune_key(mode, nrays);
  auto it = c->tuned_trace.find(key);
  if (it != c->tuned_trace.end()) {
    *tc = it->second;
    return LBVH_OK;
  }
  float best = -1.0f;
  const size_t first = c->tune_log.size();
  // one untimed run first: clocks and caches are warm for every candidate
  lbvh_status st0 = queue_trace(c, b, d_rays, nrays, mode, d_prim, d_t, d_uv, TraceConfig());
  if (st0 == LBVH_OK)
    st0 = sync_and_check(c);
  if (st0 != LBVH_OK)
    return st0;
  for (const TraceConfig& cand : trace_candidates()) {
    float ms = -1.0f;
    for (int rep = 0; rep < 3; rep++) { // the fastest of three runs
      lbvh_status st;
      const float r = time_call(c, [&] { return queue_trace(c, b, d_rays, nrays, mode, d_prim,
                                                           d_t, d_uv, cand); }, &st);
      if (st != LBVH_OK)
        return st;
      if ((st = sync_and_check(c)) != LBVH_OK)
        return st;
      if (r >= 0.0f && (ms < 0.0f || r < ms))
        ms = r;
    }
    TuneRecord r{ 0, 0, ceil_log2_u64(nrays), SortConfig(), ms, false };
    r.kind = 1;
    r.trace = cand;
    r.trace_mode = mode;
    c->tune_log.push_back(r);
    if (ms >= 0.0f && (best < 0.0f || ms < best)) {
      best = ms;
      *tc = cand;
    }
  }
  for (size_t i = first; i < c->tune_log.size(); i++)
    c->tune_log[i].picked = c->tune_log[i].trace.reorder == tc->reorder &&
                            c->tune_log[i].trace.waves_per_simd == tc->waves_per_simd;
  c->tuned_trace[key] = *tc;
  return LBVH_OK;
}

} // namespace

extern "C" {

lbvh_status
lbvh_trace_closest_device(lbvh_ctx* c, const lbvh_bvh* b, const float* d_rays, uint64_t nrays,
                          int mode, uint32_t* d_prim, float* d_t, float* d_uv)
{
  if (!c || !b || (nrays && (!d_rays || !d_prim || !d_t)))
    return fail(c, LBVH_E_INVALID, "null argument");
  if (b->prim_kind != 0)
    return fail(c, LBVH_E_INVALID, "device traversal needs a triangle tree");
  if (mode < 0 || mode > 2)
    return fail(c, LBVH_E_INVALID, "unknown intersector mode");
  if (nrays > kMaxRaysPerCall)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 rays per call");
  if (misaligned(d_rays, 16) || trace_outputs_misaligned(d_prim, d_t, d_uv))
    return fail(c, LBVH_E_INVALID,
                "device rays must be 16-byte aligned, prim and t 4-byte, uv 8-byte");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if (nrays == 0)
    return LBVH_OK;
  TraceConfig tc;
  if ((st = choose_trace(c, b, d_rays, nrays, mode, d_prim, d_t, d_uv, &tc)) != LBVH_OK)
    return st;
  return queue_trace(c, b, d_rays, nrays, mode, d_prim, d_t, d_uv, tc);
}

lbvh_status
lbvh_trace_closest(lbvh_ctx* c, const lbvh_bvh* b, const float* rays, uint64_t nrays, int mode,
                   uint32_t* prim, float* t, float* uv)
{
  if (!c || !b || (nrays && (!rays || !prim || !t)))
    return fail(c, LBVH_E_INVALID, "null argument");
  if (nrays == 0)
    return LBVH_OK;
  if (nrays > kMaxRaysPerCall)
    return fail(c, LBVH_E_INVALID, "at most 2^31 - 1 rays per call");
  lbvh_status st = set_device(c);
  if (st != LBVH_OK)
    return st;
  if ((st = ensure(c, c->io_in, nrays * 32)) != LBVH_OK)
    return st;
  if ((st = ensure(c, c->io_out, nrays * 16)) != LBVH_OK)
    return st;
  float* d_rays = static_cast<float*>(c->io_in.p);
  uint32_t* d_prim = static_cast<uint32_t*>(c->io_out.p);
  float* d_t = reinterpret_cast<float*>(d_prim + nrays);
  float* d_uv = d_t + nrays;
  HIP_TRY(c, hipMemcpyAsync(d_rays, rays, nrays * 32, hipMemcpyHostToDevice, c->stream));
  st = lbvh_trace_closest_device(c, b, d_rays, nrays, mode, d_prim, d_t, uv ? d_uv : nullptr);
  if (st != LBVH_OK)
    return st;
  HIP_TRY(c, hipMemcpyAsync(prim, d_prim, nrays * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(t, d_t, nrays * 4, hipMemcpyDeviceToHost, c->stream));
  if (uv)
    HIP_TRY(c, hipMemcpyAsync(uv, d_uv, nrays * 8, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

} // extern "C"
