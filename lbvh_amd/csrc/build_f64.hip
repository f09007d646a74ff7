// lbvh::builder<double> on gfx950 (lbvh.h:549-558 with scalar_type = double).
//
// The FP64 builder differs from builder<float> in three ways that change
// every byte of its output: codes are u64 (morton_domain<8> = 2^20 cells per
// axis, morton_encoder<8>), the arithmetic of centres, bounds and
// quantisation is FP64, and nodes are 64 bytes {aabb<double>, u64 left, u64
// right} with the leaf flag in bit 63.  The pipeline:
//   1. k_f64_bounds / k_f64_bounds_final: centroid bounds, std::min /
//      std::max semantics (lbvh.h:1282-1293, 1539-1543);
//   2. k_f64_encode: u64 codes (lbvh.h:1432-1473) and the digit histograms
//      of the radix passes the engine's sort reads (u64 keys, stable by
//      primitive id: one legal output of lbvh.h:1126-1134);
//   3. k_f64_karras: divide_node (lbvh.h:1588-1688) for every internal node
//      (builder_kernel, lbvh.h:1713-1731), plus parent links and range ends;
//   4. k_f64_fit_local / k_f64_fit_top: bottom-up box fit, the second
//      arrival at a node computes union_of(left, right) (lbvh.h:1827-1866:
//      box of the left child, then the union with the right, left operand
//      first); nodes inside a workgroup's chunk of leaves finish in LDS, the
//      few whose range crosses a chunk boundary in a second kernel.
// The input is the converter's boxes (boxes48[i] = converter(primitives[i]),
// 6 doubles, 16-byte aligned).  Traffic is HBM-bound integer/FP64 work with
// no reuse beyond the sort, so the kernels stream with grid-stride loops.
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

constexpr int kF64Block = 256;
constexpr uint32_t kF64BoundsBlocks = 1024;
constexpr uint32_t kF64EncodeBlocks = 512;
constexpr int kF64Unroll = 4;
constexpr uint64_t kLeafBit64 = 0x8000000000000000ull;

// libstdc++ std::min / std::max (the vector min/max behind union_of).
__device__ __forceinline__ double
std_min_d(double a, double b)
{
  return (b < a) ? b : a;
}
__device__ __forceinline__ double
std_max_d(double a, double b)
{
  return (a < b) ? b : a;
}

struct Box64
{
  double lo[3];
  double hi[3];
};

__device__ __forceinline__ Box64
load_box64(const double* __restrict__ boxes, uint64_t i)
{
  const double2* p = reinterpret_cast<const double2*>(boxes + 6 * i);
  double2 a = p[0], b = p[1], c = p[2];
  Box64 o;
  o.lo[0] = a.x;
  o.lo[1] = a.y;
  o.lo[2] = b.x;
  o.hi[0] = b.y;
  o.hi[1] = c.x;
  o.hi[2] = c.y;
  return o;
}

// center_of (lbvh.h:1164-1167): (min + max) * 0.5 per component.
__device__ __forceinline__ void
centre64(const Box64& b, double c[3])
{
#pragma unroll
  for (int k = 0; k < 3; k++)
    c[k] = (b.lo[k] + b.hi[k]) * 0.5;
}

// Workgroup-wide std::min / std::max of lo[3] / hi[3] into s[6]; accumulators
// start at +-inf and a NaN centre never replaces them, so the combination
// order does not change the result (signed zeros aside, which quantise alike).
__device__ __forceinline__ void
block_bounds(double lo[3], double hi[3], double* s_part)
{
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = std_min_d(lo[k], __shfl_xor(lo[k], off, kWave));
      hi[k] = std_max_d(hi[k], __shfl_xor(hi[k], off, kWave));
    }
  }
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_part[wave * 6 + k] = lo[k];
      s_part[wave * 6 + 3 + k] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kF64Block / kWave; w++) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        s_part[k] = std_min_d(s_part[k], s_part[w * 6 + k]);
        s_part[3 + k] = std_max_d(s_part[3 + k], s_part[w * 6 + 3 + k]);
      }
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kF64Block) void
k_f64_bounds(const double* __restrict__ boxes, uint64_t n, double* __restrict__ partials)
{
  __shared__ double s_part[6 * (kF64Block / kWave)];
  double lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = __builtin_inf();
    hi[k] = -__builtin_inf();
  }
  const uint64_t stride = uint64_t(gridDim.x) * kF64Block;
  for (uint64_t i0 = uint64_t(blockIdx.x) * kF64Block + threadIdx.x; i0 < n;
       i0 += stride * kF64Unroll) {
    // all loads in flight together (past the end: the last box again, which
    // changes no minimum or maximum)
    Box64 b[kF64Unroll];
#pragma unroll
    for (int u = 0; u < kF64Unroll; u++) {
      const uint64_t i = i0 + u * stride;
      b[u] = load_box64(boxes, i < n ? i : n - 1);
    }
#pragma unroll
    for (int u = 0; u < kF64Unroll; u++) {
      double c[3];
      centre64(b[u], c);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = std_min_d(lo[k], c[k]);
        hi[k] = std_max_d(hi[k], c[k]);
      }
    }
  }
  block_bounds(lo, hi, s_part);
  if (threadIdx.x < 6)
    partials[size_t(blockIdx.x) * 6 + threadIdx.x] = s_part[threadIdx.x];
}

__global__ __launch_bounds__(kF64Block) void
k_f64_bounds_final(const double* __restrict__ partials, uint32_t nparts,
                   double* __restrict__ bounds)
{
  __shared__ double s_part[6 * (kF64Block / kWave)];
  double lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = __builtin_inf();
    hi[k] = -__builtin_inf();
  }
  for (uint32_t p = threadIdx.x; p < nparts; p += kF64Block) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = std_min_d(lo[k], partials[size_t(p) * 6 + k]);
      hi[k] = std_max_d(hi[k], partials[size_t(p) * 6 + 3 + k]);
    }
  }
  block_bounds(lo, hi, s_part);
  if (threadIdx.x < 6)
    bounds[threadIdx.x] = s_part[threadIdx.x];
}

// morton_encoder<8>::expand (lbvh.h:1377-1385): 20 bits to every third bit.
__device__ __forceinline__ uint64_t
expand20_d(uint64_t n)
{
  n = (n | n << 32) & 0x001f00000000ffffull;
  n = (n | n << 16) & 0x001f0000ff0000ffull;
  n = (n | n << 8) & 0x100f00f00f00f00full;
  n = (n | n << 4) & 0x10c30c30c30c30c3ull;
  n = (n | n << 2) & 0x1249249249249249ull;
  return n;
}

// Codes (lbvh.h:1432-1473 in FP64: scale = 2^20 / (extent + 1), q = (c -
// min) * scale truncated to u64) and the digit histograms of passes
// first_pass .. passes - 1, counted in LDS and added to hist[(p - first_pass)
// * radix + d] (zero before).
__global__ __launch_bounds__(kF64Block) void
k_f64_encode(const double* __restrict__ boxes, uint64_t n, const double* __restrict__ bounds,
             uint64_t* __restrict__ codes, uint32_t* __restrict__ hist, int digit_bits,
             int first_pass, int passes)
{
  extern __shared__ uint32_t s_hist[];
  const int radix = 1 << digit_bits;
  const int words = radix * (passes - first_pass);
  for (int j = threadIdx.x; j < words; j += kF64Block)
    s_hist[j] = 0u;
  double bmin[3], scale[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bmin[k] = bounds[k];
    double ext = bounds[3 + k] - bmin[k];
    scale[k] = 1048576.0 / (ext + 1.0);
  }
  __syncthreads();
  const uint64_t mask = uint64_t(radix - 1);
  const uint64_t stride = uint64_t(gridDim.x) * kF64Block;
  for (uint64_t i0 = uint64_t(blockIdx.x) * kF64Block + threadIdx.x; i0 < n;
       i0 += stride * kF64Unroll) {
    Box64 b[kF64Unroll];
#pragma unroll
    for (int u = 0; u < kF64Unroll; u++) {
      const uint64_t i = i0 + u * stride;
      b[u] = load_box64(boxes, i < n ? i : n - 1);
    }
#pragma unroll
    for (int u = 0; u < kF64Unroll; u++) {
      const uint64_t i = i0 + u * stride;
      if (i >= n)
        continue;
      double c[3];
      centre64(b[u], c);
      uint64_t q[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double d = c[k] - bmin[k];
        q[k] = uint64_t(d * scale[k]);
      }
      const uint64_t code =
        (expand20_d(q[0]) << 2) | (expand20_d(q[1]) << 1) | expand20_d(q[2]);
      codes[i] = code;
      for (int p = first_pass; p < passes; p++)
        atomicAdd(&s_hist[(p - first_pass) * radix + int((code >> (p * digit_bits)) & mask)],
                  1u);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < words; j += kF64Block) {
    const uint32_t v = s_hist[j];
    if (v)
      atomicAdd(&hist[j], v);
  }
}

// Sorted codes around a workgroup's nodes, kept in LDS: positions [lo, hi)
// read from LDS, the rest (far probes of large ranges) from memory.
struct CodeWindow
{
  const uint64_t* __restrict__ g;
  const uint64_t* s;
  int64_t lo, hi;
  __device__ __forceinline__ uint64_t at(int64_t k) const
  {
    return (k >= lo && k < hi) ? s[k - lo] : g[k];
  }
};

// calc_delta (lbvh.h:1602-1621) for u64 codes: -1 out of range, clz of the
// code difference, or clz(j ^ k) + 64 for equal codes.  code_j is the code
// at j (the node's own position).
__device__ __forceinline__ int64_t
delta64(const CodeWindow& c, int64_t n, int64_t j, uint64_t code_j, int64_t k)
{
  if (k < 0 || k >= n)
    return -1;
  const uint64_t b = c.at(k);
  if (code_j == b)
    return int64_t(__clzll((long long)(uint64_t(j) ^ uint64_t(k)))) + 64;
  return __clzll((long long)(code_j ^ b));
}

// Nodes per Karras workgroup and the codes kept in LDS on either side.
constexpr int kKarrasBlock = 1024;
constexpr int kKarrasMargin = 1024;

// divide_node (lbvh.h:1588-1688) for internal node i, int64 offsets as the
// reference's int_type for u64 codes; writes left / right (leaf flag bit 63),
// the other end j of the node's leaf range [min(i, j), max(i, j)], and the
// parent link of both children (internal child c at parent[c], leaf k at
// parent[n - 1 + k]; bit 31 set for a right child).
__global__ __launch_bounds__(kKarrasBlock) void
k_f64_karras(const uint64_t* __restrict__ codes, uint64_t n_, const uint32_t* __restrict__ ids,
             uint64_t* __restrict__ node_words, uint32_t* __restrict__ parent,
             uint32_t* __restrict__ other_end)
{
  __shared__ uint64_t s_codes[kKarrasBlock + 2 * kKarrasMargin];
  const int64_t n = int64_t(n_);
  const int64_t i0 = int64_t(blockIdx.x) * kKarrasBlock;
  CodeWindow cw;
  cw.g = codes;
  cw.s = s_codes;
  cw.lo = i0 > kKarrasMargin ? i0 - kKarrasMargin : 0;
  cw.hi = i0 + kKarrasBlock + kKarrasMargin < n ? i0 + kKarrasBlock + kKarrasMargin : n;
  for (int64_t k = cw.lo + threadIdx.x; k < cw.hi; k += kKarrasBlock)
    s_codes[k - cw.lo] = codes[k];
  __syncthreads();
  const int64_t i = i0 + threadIdx.x;
  if (i >= n - 1)
    return;
  const uint64_t ci = cw.at(i);
  const int64_t d = (delta64(cw, n, i, ci, i + 1) - delta64(cw, n, i, ci, i - 1)) < 0 ? -1 : 1;
  const int64_t delta_min = delta64(cw, n, i, ci, i - d);
  int64_t l_max = 128;
  for (;;) {
    const int64_t k = i + l_max * d;
    if (k >= n)
      break;
    if (!(delta64(cw, n, i, ci, k) > delta_min))
      break;
    l_max *= 4;
  }
  // div = 2^sh: the reference's l_max / div and ceil_div(l, div) on
  // non-negative values are shifts
  int64_t l = 0;
  for (int sh = 1;; sh++) {
    const int64_t t = l_max >> sh;
    if (delta64(cw, n, i, ci, i + (l + t) * d) > delta_min)
      l += t;
    if (t == 1)
      break;
  }
  const int64_t j = i + l * d;
  const int64_t delta_node = delta64(cw, n, i, ci, j);
  int64_t s = 0;
  for (int sh = 1;; sh++) {
    const int64_t div = int64_t(1) << sh;
    const int64_t t = (l + (div - 1)) >> sh;
    if (delta64(cw, n, i, ci, i + (s + t) * d) > delta_node)
      s += t;
    if (div >= l)
      break;
  }
  const int64_t g = i + s * d + (d < 0 ? d : 0);
  const int64_t mn = i < j ? i : j;
  const int64_t mx = j > i ? j : i;
  const bool l_leaf = mn == g, r_leaf = mx == g + 1;
  uint64_t left = uint64_t(g), right = uint64_t(g + 1);
  // a leaf names its sorted position, or (ids != nullptr) the primitive there
  if (l_leaf)
    left = (ids ? uint64_t(ids[g]) : left) | kLeafBit64;
  if (r_leaf)
    right = (ids ? uint64_t(ids[g + 1]) : right) | kLeafBit64;
  node_words[size_t(i) * 8 + 6] = left;
  node_words[size_t(i) * 8 + 7] = right;
  other_end[i] = uint32_t(j);
  parent[l_leaf ? size_t(n - 1 + g) : size_t(g)] = uint32_t(i);
  parent[r_leaf ? size_t(n - 1 + g + 1) : size_t(g + 1)] = uint32_t(i) | 0x80000000u;
}

__device__ __forceinline__ Box64
box_union64(const Box64& l, const Box64& r)
{
  Box64 o;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    o.lo[k] = std_min_d(l.lo[k], r.lo[k]);
    o.hi[k] = std_max_d(l.hi[k], r.hi[k]);
  }
  return o;
}

__device__ __forceinline__ void
store_box64(double* p, const Box64& b)
{
  double2* q = reinterpret_cast<double2*>(p);
  q[0] = make_double2(b.lo[0], b.lo[1]);
  q[1] = make_double2(b.lo[2], b.hi[0]);
  q[2] = make_double2(b.hi[1], b.hi[2]);
}

// Leaves per fit workgroup (one thread each).
constexpr int kF64FitLeaves = 512;
constexpr uint32_t kQueueLeaf = 0x80000000u;

// Bottom-up fit, part 1: one thread per sorted leaf position.  A node whose
// leaf range lies inside the workgroup's chunk of kF64FitLeaves positions is
// finished in LDS: each arriving child leaves its box in the node's left or
// right LDS slot and counts its arrival; the second arrival takes the union
// (left operand first, union_of of lbvh.h:1208-1213) and climbs on.  A child
// whose parent's range crosses the chunk boundary stops there: its box goes
// to `stage` and the child is queued for part 2 in the chunk's own queue
// segment (2 kF64FitLeaves entries: queued nodes are leaves or internal
// nodes numbered inside the chunk), its count in queue_count[chunk].
__device__ __forceinline__ void
climb_local(const double* __restrict__ boxes, uint64_t n, const uint32_t* __restrict__ ids,
            uint64_t* __restrict__ node_words, uint32_t leaf_link, double* s_box,
            uint32_t* s_flag, const uint32_t* s_oe, const uint32_t* s_par, uint64_t first,
            uint64_t k, double* __restrict__ stage, uint32_t* __restrict__ my_queue,
            uint32_t* s_queued)
{
  const uint64_t last = (first + kF64FitLeaves < n ? first + kF64FitLeaves : n) - 1;
  Box64 box = load_box64(boxes, ids ? uint64_t(ids[k]) : k);
  uint64_t child = kLeafBit64 | k;
  uint32_t link = leaf_link;
  for (;;) {
    const uint32_t p = link & 0x7fffffffu;
    const bool right = (link >> 31) != 0;
    const bool in_chunk = p >= first && p <= last; // p <= n - 2 always
    const uint32_t oe = in_chunk ? s_oe[p - first] : 0u;
    const bool local = in_chunk && (p < oe ? p : oe) >= first && (p > oe ? p : oe) <= last;
    if (!local) {
      if (!(child & kLeafBit64))
        store_box64(stage + child * 6, box);
      const uint32_t e = (child & kLeafBit64) ? (uint32_t(k) | kQueueLeaf) : uint32_t(child);
      my_queue[atomicAdd(s_queued, 1u)] = e;
      return;
    }
    const uint32_t slot = p - uint32_t(first);
    double* mine = s_box + (slot * 2 + (right ? 1 : 0)) * 6;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      mine[c] = box.lo[c];
      mine[3 + c] = box.hi[c];
    }
    if (__hip_atomic_fetch_add(&s_flag[slot], 1u, __ATOMIC_ACQ_REL,
                               __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
      return;
    const double* other = s_box + (slot * 2 + (right ? 0 : 1)) * 6;
    Box64 ob;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      ob.lo[c] = other[c];
      ob.hi[c] = other[3 + c];
    }
    box = right ? box_union64(ob, box) : box_union64(box, ob);
    store_box64(reinterpret_cast<double*>(node_words + size_t(p) * 8), box);
    child = p;
    if (p == 0)
      return;
    link = s_par[p - first];
  }
}

__global__ __launch_bounds__(kF64FitLeaves) void
k_f64_fit_local(const double* __restrict__ boxes, uint64_t n, const uint32_t* __restrict__ ids,
                uint64_t* __restrict__ node_words, const uint32_t* __restrict__ parent,
                const uint32_t* __restrict__ other_end, double* __restrict__ stage,
                uint32_t* __restrict__ queue, uint32_t* __restrict__ queue_count)
{
  __shared__ double s_box[kF64FitLeaves * 2 * 6];
  __shared__ uint32_t s_flag[kF64FitLeaves];
  // range ends and parent links of the internal nodes numbered inside the
  // chunk (every node finished in LDS is one of them)
  __shared__ uint32_t s_oe[kF64FitLeaves];
  __shared__ uint32_t s_par[kF64FitLeaves];
  __shared__ uint32_t s_queued;
  const uint64_t first = uint64_t(blockIdx.x) * kF64FitLeaves;
  const uint64_t k = first + threadIdx.x;
  s_flag[threadIdx.x] = 0u;
  if (threadIdx.x == 0)
    s_queued = 0u;
  if (k < n - 1) {
    s_oe[threadIdx.x] = other_end[k];
    s_par[threadIdx.x] = parent[k];
  }
  __syncthreads();
  if (k < n)
    climb_local(boxes, n, ids, node_words, parent[n - 1 + k], s_box, s_flag, s_oe, s_par,
                first, k, stage, queue + size_t(blockIdx.x) * (2 * kF64FitLeaves), &s_queued);
  __syncthreads();
  if (threadIdx.x == 0)
    queue_count[blockIdx.x] = s_queued;
}

// Bottom-up fit, part 2: the nodes whose range crosses a chunk boundary
// (all ancestors of such a node do too).  One wave per chunk takes the
// chunk's queued children; each climbs with a device-scope arrival count
// per node: the second arrival reads the sibling's box (input box for a
// leaf, `stage` otherwise), writes the node's box, publishes it to `stage`
// with device-coherent stores and drains them before counting its arrival
// at the next node.  flags[] must be zero.
__global__ __launch_bounds__(kWave) void
k_f64_fit_top(const double* __restrict__ boxes, uint64_t n, const uint32_t* __restrict__ ids,
              uint64_t* __restrict__ node_words, const uint32_t* __restrict__ parent,
              uint32_t* flags, uint64_t* stage, const uint32_t* __restrict__ queue,
              const uint32_t* __restrict__ queue_count)
{
  const uint32_t count = queue_count[blockIdx.x];
  const uint32_t* my_queue = queue + size_t(blockIdx.x) * (2 * kF64FitLeaves);
  for (uint32_t q = threadIdx.x; q < count; q += kWave) {
    const uint32_t e = my_queue[q];
    const uint32_t c0 = e & ~kQueueLeaf;
    Box64 box;
    uint32_t link;
    if (e & kQueueLeaf) {
      box = load_box64(boxes, ids ? uint64_t(ids[c0]) : uint64_t(c0));
      link = parent[n - 1 + c0];
    } else {
      box = load_box64(reinterpret_cast<const double*>(stage), c0);
      link = parent[c0];
    }
    for (;;) {
      const uint32_t p = link & 0x7fffffffu;
      const bool right = (link >> 31) != 0;
      // the sibling and the next parent link do not depend on the arrival
      // count: their loads complete under the drain below
      const uint64_t sib = node_words[size_t(p) * 8 + (right ? 6 : 7)];
      const uint32_t next = parent[p];
      vmem_drain();
      if (atomic_add_agent(&flags[p], 1u) == 0u)
        break;
      Box64 ob;
      if (sib & kLeafBit64) {
        ob = load_box64(boxes, sib & ~kLeafBit64);
      } else {
        const uint64_t* st = stage + sib * 6;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          ob.lo[c] = __longlong_as_double((long long)coherent_load_u64(st + c));
          ob.hi[c] = __longlong_as_double((long long)coherent_load_u64(st + 3 + c));
        }
      }
      box = right ? box_union64(ob, box) : box_union64(box, ob);
      store_box64(reinterpret_cast<double*>(node_words + size_t(p) * 8), box);
      if (p == 0)
        break;
      uint64_t* st = stage + size_t(p) * 6;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        coherent_store_u64(st + c, (uint64_t)__double_as_longlong(box.lo[c]));
        coherent_store_u64(st + 3 + c, (uint64_t)__double_as_longlong(box.hi[c]));
      }
      link = next;
    }
  }
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

uint32_t
f64_bounds_blocks(uint64_t n)
{
  uint64_t b = (n + kF64Block * kF64Unroll - 1) / (kF64Block * kF64Unroll);
  return uint32_t(b < kF64BoundsBlocks ? (b ? b : 1) : kF64BoundsBlocks);
}

hipError_t
launch_f64_bounds(const double* boxes, uint64_t n, double* partials, double* bounds, hipStream_t s)
{
  const uint32_t g = f64_bounds_blocks(n);
  hipLaunchKernelGGL(k_f64_bounds, dim3(g), dim3(kF64Block), 0, s, boxes, n, partials);
  hipLaunchKernelGGL(k_f64_bounds_final, dim3(1), dim3(kF64Block), 0, s,
                     static_cast<const double*>(partials), g, bounds);
  return hipGetLastError();
}

hipError_t
launch_f64_encode(const double* boxes, uint64_t n, const double* bounds, uint64_t* codes,
                  uint32_t* hist, int digit_bits, int first_pass, int passes, hipStream_t s)
{
  const size_t lds = (size_t(4) << digit_bits) * size_t(passes - first_pass);
  if (first_pass < 0 || first_pass >= passes || lds > 64 * 1024)
    return hipErrorInvalidValue;
  // few workgroups: each flushes up to passes x radix histogram words
  uint64_t b = (n + kF64Block * kF64Unroll - 1) / (kF64Block * kF64Unroll);
  const uint32_t g = uint32_t(b < kF64EncodeBlocks ? (b ? b : 1) : kF64EncodeBlocks);
  // the kernel counts pass p at (p - first_pass) * radix of its pointer
  hipLaunchKernelGGL(k_f64_encode, dim3(g), dim3(kF64Block), lds, s, boxes, n, bounds, codes,
                     hist + size_t(first_pass) * (size_t(1) << digit_bits), digit_bits,
                     first_pass, passes);
  return hipGetLastError();
}

// Scratch of the FP64 tree, in order: parent links [2n - 1], range ends
// [n - 1], fit flags [n - 1], queue counts [chunks], fit queues [chunks * 2
// kF64FitLeaves] (u32), then 16-byte aligned staging boxes [n - 1] of 48
// bytes.
struct F64Scratch
{
  uint32_t *parent, *other_end, *flags, *count, *queue;
  double* stage;
};

static uint64_t
f64_fit_chunks(uint64_t n)
{
  return (n + kF64FitLeaves - 1) / kF64FitLeaves;
}

static uint64_t
f64_scratch_words(uint64_t n)
{
  return (2 * n - 1) + 2 * (n - 1) + f64_fit_chunks(n) * (1 + 2 * kF64FitLeaves);
}

static F64Scratch
f64_scratch(void* base, uint64_t n)
{
  F64Scratch f;
  f.parent = static_cast<uint32_t*>(base);
  f.other_end = f.parent + (2 * n - 1);
  f.flags = f.other_end + (n - 1);
  f.count = f.flags + (n - 1);
  f.queue = f.count + f64_fit_chunks(n);
  f.stage = reinterpret_cast<double*>(static_cast<char*>(base) +
                                      ((4 * f64_scratch_words(n) + 15) & ~uint64_t(15)));
  return f;
}

uint64_t
f64_tree_scratch_bytes(uint64_t n)
{
  return ((4 * f64_scratch_words(n) + 15) & ~uint64_t(15)) + 48 * (n - 1);
}

hipError_t
launch_f64_tree(const uint64_t* sorted, const uint32_t* ids, const double* boxes, uint64_t n,
                void* nodes, void* scratch, hipStream_t s)
{
  if (n < 2)
    return hipErrorInvalidValue;
  uint64_t* w = static_cast<uint64_t*>(nodes);
  const F64Scratch f = f64_scratch(scratch, n);
  const uint64_t gi = (n - 1 + kKarrasBlock - 1) / kKarrasBlock;
  hipLaunchKernelGGL(k_f64_karras, dim3(uint32_t(gi)), dim3(kKarrasBlock), 0, s, sorted, n, ids,
                     w, f.parent, f.other_end);
  hipError_t e = launch_fill_u32(f.flags, n - 1, 0u, s);
  if (e != hipSuccess)
    return e;
  const uint32_t chunks = uint32_t(f64_fit_chunks(n));
  hipLaunchKernelGGL(k_f64_fit_local, dim3(chunks), dim3(kF64FitLeaves), 0, s, boxes, n, ids, w,
                     static_cast<const uint32_t*>(f.parent),
                     static_cast<const uint32_t*>(f.other_end), f.stage, f.queue, f.count);
  hipLaunchKernelGGL(k_f64_fit_top, dim3(chunks), dim3(kWave), 0, s, boxes, n, ids, w,
                     static_cast<const uint32_t*>(f.parent), f.flags,
                     reinterpret_cast<uint64_t*>(f.stage), static_cast<const uint32_t*>(f.queue),
                     static_cast<const uint32_t*>(f.count));
  return hipGetLastError();
}

} // namespace lbvh_eng
