// Host runtime internals shared by the C ABI translation units: the context
// and tree handles, the context's small-word layout, error helpers and the
// grow-only workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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
constexpr uint32_t kSmallBoundsTicket = 0; // self-resetting
constexpr uint32_t kSmallErr = 1;          // cleared by the host after reading
constexpr uint32_t kSmallTickets = 2;      // kMaxSortPasses words
constexpr uint32_t kSmallTreeletCount = kSmallTickets + kMaxSortPasses;
constexpr uint32_t kSmallTreeletQueue = kSmallTreeletCount + 1;
constexpr uint32_t kSmallHist = kSmallTreeletQueue + 1;

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
  // workspace
  lbvh_rt::DevBuf cent, keys_a, keys_b, ids_a, ids_b, partials, part_hist, small, status, counts,
    treelets, io_in, io_out;
};

namespace lbvh_rt {

lbvh_status fail(lbvh_ctx* c, lbvh_status s, const std::string& msg);
lbvh_status hip_fail(lbvh_ctx* c, hipError_t e, const char* what);
lbvh_status ensure(lbvh_ctx* c, DevBuf& b, uint64_t bytes, bool zero_new = false);
void dev_free(void* p);
lbvh_status set_device(lbvh_ctx* c);
void begin_timed(lbvh_ctx* c, int kind);
void record(lbvh_ctx* c, int idx, bool final = false);
SortConfig tune_sort(uint64_t n, uint32_t code_bits, const lbvh_build_opts* o);
bool valid_sort_config(const SortConfig& c);
lbvh_status sync_and_check(lbvh_ctx* c);

} // namespace lbvh_rt

#define HIP_TRY(ctx, expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return lbvh_rt::hip_fail((ctx), e_, #expr);                              \
  } while (0)
