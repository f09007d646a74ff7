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
//      of every radix pass, for the engine's radix sort (u64 keys, stable
//      by primitive id: one legal output of lbvh.h:1126-1134);
//   3. k_f64_karras: divide_node (lbvh.h:1588-1688) for every internal node
//      (builder_kernel, lbvh.h:1713-1731), plus parent links;
//   4. k_f64_fit: bottom-up box fit, one thread per leaf, the second arrival
//      at a node computes union_of(left, right) (lbvh.h:1827-1866: box of the
//      left child, then the union with the right, left operand first).
// The input is the converter's boxes (boxes48[i] = converter(primitives[i]),
// 6 doubles, 16-byte aligned).  Traffic is HBM-bound integer/FP64 work with
// no reuse beyond the sort, so the kernels stream with grid-stride loops.
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

constexpr int kF64Block = 256;
constexpr uint32_t kF64BoundsBlocks = 1024;
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
  for (uint64_t i = uint64_t(blockIdx.x) * kF64Block + threadIdx.x; i < n; i += stride) {
    double c[3];
    centre64(load_box64(boxes, i), c);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = std_min_d(lo[k], c[k]);
      hi[k] = std_max_d(hi[k], c[k]);
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
// min) * scale truncated to u64) and the digit histograms of every pass,
// counted in LDS and added to hist[p * radix + d] (zero before).
__global__ __launch_bounds__(kF64Block) void
k_f64_encode(const double* __restrict__ boxes, uint64_t n, const double* __restrict__ bounds,
             uint64_t* __restrict__ codes, uint32_t* __restrict__ hist, int digit_bits,
             int passes)
{
  extern __shared__ uint32_t s_hist[];
  const int radix = 1 << digit_bits;
  const int words = radix * passes;
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
  for (uint64_t i = uint64_t(blockIdx.x) * kF64Block + threadIdx.x; i < n; i += stride) {
    double c[3];
    centre64(load_box64(boxes, i), c);
    uint64_t q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double d = c[k] - bmin[k];
      q[k] = uint64_t(d * scale[k]);
    }
    const uint64_t code = (expand20_d(q[0]) << 2) | (expand20_d(q[1]) << 1) | expand20_d(q[2]);
    codes[i] = code;
    for (int p = 0; p < passes; p++)
      atomicAdd(&s_hist[p * radix + int((code >> (p * digit_bits)) & mask)], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < words; j += kF64Block) {
    const uint32_t v = s_hist[j];
    if (v)
      atomicAdd(&hist[j], v);
  }
}

// calc_delta (lbvh.h:1602-1621) for u64 codes: -1 out of range, clz of the
// code difference, or clz(j ^ k) + 64 for equal codes.
__device__ __forceinline__ int64_t
delta64(const uint64_t* __restrict__ c, int64_t n, int64_t j, int64_t k)
{
  if (k < 0 || k >= n)
    return -1;
  const uint64_t a = c[j], b = c[k];
  if (a == b)
    return int64_t(__clzll((long long)(uint64_t(j) ^ uint64_t(k)))) + 64;
  return __clzll((long long)(a ^ b));
}

// divide_node (lbvh.h:1588-1688) for internal node i, int64 offsets as the
// reference's int_type for u64 codes; writes left / right (leaf flag bit 63)
// and the parent of both children: internal child c at parent[c], leaf k at
// parent[n - 1 + k].
__global__ __launch_bounds__(kF64Block) void
k_f64_karras(const uint64_t* __restrict__ codes, uint64_t n_, const uint32_t* __restrict__ ids,
             uint64_t* __restrict__ node_words, uint32_t* __restrict__ parent)
{
  const int64_t n = int64_t(n_);
  const int64_t i = int64_t(blockIdx.x) * kF64Block + threadIdx.x;
  if (i >= n - 1)
    return;
  const int64_t d = (delta64(codes, n, i, i + 1) - delta64(codes, n, i, i - 1)) < 0 ? -1 : 1;
  const int64_t delta_min = delta64(codes, n, i, i - d);
  int64_t l_max = 128;
  for (;;) {
    const int64_t k = i + l_max * d;
    if (k >= n)
      break;
    if (!(delta64(codes, n, i, k) > delta_min))
      break;
    l_max *= 4;
  }
  int64_t l = 0;
  for (int64_t div = 2;; div *= 2) {
    const int64_t t = l_max / div;
    if (delta64(codes, n, i, i + (l + t) * d) > delta_min)
      l += t;
    if (t == 1)
      break;
  }
  const int64_t j = i + l * d;
  const int64_t delta_node = delta64(codes, n, i, j);
  int64_t s = 0;
  for (int64_t div = 2;; div *= 2) {
    const int64_t t = (l + (div - 1)) / div;
    if (delta64(codes, n, i, i + (s + t) * d) > delta_node)
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
  parent[l_leaf ? size_t(n - 1 + g) : size_t(g)] = uint32_t(i);
  parent[r_leaf ? size_t(n - 1 + g + 1) : size_t(g + 1)] = uint32_t(i);
}

__device__ __forceinline__ Box64
child_box(const double* __restrict__ boxes, const uint64_t* node_words, uint64_t ch)
{
  if (ch & kLeafBit64)
    return load_box64(boxes, ch & ~kLeafBit64);
  // written by another thread of this launch: read past the non-coherent
  // caches (the writer stored the same way)
  Box64 o;
  const uint64_t* w = node_words + ch * 8;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    o.lo[k] = __longlong_as_double((long long)coherent_load_u64(w + k));
    o.hi[k] = __longlong_as_double((long long)coherent_load_u64(w + 3 + k));
  }
  return o;
}

// Bottom-up fit: leaf thread k climbs from its parent; at every node the
// first arrival stops and the second (both children final) writes the box
// and climbs on.  flags[] must be zero.  Box stores are device-coherent and
// drained before the arrival count of the next node is taken.
__global__ __launch_bounds__(kF64Block) void
k_f64_fit(const double* __restrict__ boxes, uint64_t n, uint64_t* node_words,
          const uint32_t* __restrict__ parent, uint32_t* flags)
{
  const uint64_t k = uint64_t(blockIdx.x) * kF64Block + threadIdx.x;
  if (k >= n)
    return;
  uint32_t p = parent[n - 1 + k];
  for (;;) {
    if (atomic_add_agent(&flags[p], 1u) == 0u)
      return;
    const uint64_t left = node_words[size_t(p) * 8 + 6];
    const uint64_t right = node_words[size_t(p) * 8 + 7];
    const Box64 a = child_box(boxes, node_words, left);
    const Box64 b = child_box(boxes, node_words, right);
    uint64_t* w = node_words + size_t(p) * 8;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      coherent_store_u64(w + c, (uint64_t)__double_as_longlong(std_min_d(a.lo[c], b.lo[c])));
      coherent_store_u64(w + 3 + c, (uint64_t)__double_as_longlong(std_max_d(a.hi[c], b.hi[c])));
    }
    if (p == 0)
      return;
    vmem_drain();
    p = parent[p];
  }
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

uint32_t
f64_bounds_blocks(uint64_t n)
{
  uint64_t b = (n + kF64Block - 1) / kF64Block;
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
                  uint32_t* hist, int digit_bits, int passes, hipStream_t s)
{
  const size_t lds = size_t(4) << digit_bits;
  if (lds * passes > 64 * 1024)
    return hipErrorInvalidValue;
  uint64_t b = (n + kF64Block - 1) / kF64Block;
  const uint32_t g = uint32_t(b < 2048 ? (b ? b : 1) : 2048);
  hipLaunchKernelGGL(k_f64_encode, dim3(g), dim3(kF64Block), lds * passes, s, boxes, n, bounds,
                     codes, hist, digit_bits, passes);
  return hipGetLastError();
}

hipError_t
launch_f64_tree(const uint64_t* sorted, const uint32_t* ids, const double* boxes, uint64_t n,
                void* nodes, uint32_t* parent, uint32_t* flags, hipStream_t s)
{
  if (n < 2)
    return hipErrorInvalidValue;
  uint64_t* w = static_cast<uint64_t*>(nodes);
  const uint64_t gi = (n - 1 + kF64Block - 1) / kF64Block;
  hipLaunchKernelGGL(k_f64_karras, dim3(uint32_t(gi)), dim3(kF64Block), 0, s, sorted, n, ids, w,
                     parent);
  hipError_t e = launch_fill_u32(flags, n - 1, 0u, s);
  if (e != hipSuccess)
    return e;
  const uint64_t gl = (n + kF64Block - 1) / kF64Block;
  hipLaunchKernelGGL(k_f64_fit, dim3(uint32_t(gl)), dim3(kF64Block), 0, s, boxes, n, w,
                     static_cast<const uint32_t*>(parent), flags);
  return hipGetLastError();
}

} // namespace lbvh_eng
