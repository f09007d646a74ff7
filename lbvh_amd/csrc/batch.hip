// gfx950 kernels of the segmented (many-scene) build: per-scene centroid
// bounds and per-scene 30-bit Morton codes keyed by scene.  The tree, fit and
// traversal kernels take the scene layout directly (see Scenes).
//
// The reference builds one scene per call (builder::operator(),
// lbvh.h:1793-1822); a batch applies the same arithmetic to every scene with
// that scene's own bounds: centroid_bounds_kernel (lbvh.h:1248-1305) per
// scene, morton_curve_kernel with morton_domain<4> / morton_encoder<4>
// (lbvh.h:1421-1475, 1338-1360) against the scene's bounds.
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

constexpr int kSegBlock = 256;
constexpr int kSegMaxBlocks = 4096;
constexpr int kSegUnroll = 8;
constexpr int kSegHistMaxWords = 8192;

// Order-preserving map of floats to signed integers (x < y as floats iff
// key(x) < key(y), with -0 below +0), so that per-scene min/max can be
// accumulated with integer atomics.  min/max are exact, so the result is the
// reference's serial union for finite centres (up to the sign of a zero
// bound, which the encode cannot observe: c - (+-0) and ext + 1 are the same
// for either sign).
__device__ __forceinline__ int32_t
order_key(float f)
{
  const int32_t b = __float_as_int(f);
  return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ __forceinline__ float
order_value(int32_t k)
{
  return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff));
}

// Each workgroup owns a contiguous range of primitives; scene lookups advance
// monotonically from the range's first scene.
struct BlockRange
{
  uint64_t begin, end;
};

__device__ __forceinline__ BlockRange
block_range(uint64_t n)
{
  const uint64_t per = ((n + gridDim.x - 1) / gridDim.x + kSegBlock - 1) / kSegBlock * kSegBlock;
  BlockRange r;
  r.begin = uint64_t(blockIdx.x) * per;
  r.end = r.begin + per < n ? r.begin + per : n;
  if (r.begin > n)
    r.begin = n;
  return r;
}

__device__ __forceinline__ void
wave_reduce_min_max(float lo[3], float hi[3])
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float ol = __shfl_down(lo[k], o, kWave);
      const float oh = __shfl_down(hi[k], o, kWave);
      lo[k] = std_min(lo[k], ol);
      hi[k] = std_max(hi[k], oh);
    }
  }
}

__device__ __forceinline__ uint32_t
advance_scene(const uint64_t* __restrict__ off, uint32_t nseg, uint32_t s, uint64_t i)
{
  while (s + 1 < nseg && off[s + 1] <= i)
    s++;
  return s;
}

__global__ __launch_bounds__(kSegBlock) void
k_scene_bounds_init(int32_t* __restrict__ keys, uint32_t nseg)
{
  const uint32_t i = blockIdx.x * kSegBlock + threadIdx.x;
  if (i >= 6 * nseg)
    return;
  keys[i] = (i % 6) < 3 ? order_key(INFINITY) : order_key(-INFINITY);
}

__device__ __forceinline__ void
flush_box(int32_t* __restrict__ keys, uint32_t s, const float lo[3], const float hi[3])
{
#pragma unroll
  for (int k = 0; k < 3; k++) {
    atomicMin(&keys[6 * s + k], order_key(lo[k]));
    atomicMax(&keys[6 * s + 3 + k], order_key(hi[k]));
  }
}

// KIND: kPrimTriangle or kPrimAabb (a template argument, so that a thread's
// kSegUnroll loads issue together).
template<int KIND>
__global__ __launch_bounds__(kSegBlock) void
k_scene_bounds(const float* __restrict__ prims, uint64_t n,
               const uint64_t* __restrict__ off, uint32_t nseg, float* __restrict__ cx,
               float* __restrict__ cy, float* __restrict__ cz, int32_t* __restrict__ keys)
{
  __shared__ uint32_t s_first;
  const BlockRange r = block_range(n);
  if (r.begin >= r.end)
    return;
  if (threadIdx.x == 0)
    s_first = find_scene(off, nseg, int64_t(r.begin), 1, false);
  __syncthreads();
  const int lane = __lane_id();
  const int wv = threadIdx.x / kWave;
  uint32_t s = s_first;
  // the current run: a scene every lane of the wave was in, with per-lane
  // partial boxes
  uint32_t run = 0xffffffffu;
  float rlo[3] = { INFINITY, INFINITY, INFINITY };
  float rhi[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (uint64_t t0 = r.begin; t0 < r.end; t0 += uint64_t(kSegBlock) * kSegUnroll) {
    // unconditional loads (past the end: the range's last primitive, unused)
    Box b[kSegUnroll];
#pragma unroll
    for (int u = 0; u < kSegUnroll; u++) {
      const uint64_t i = t0 + uint64_t(u) * kSegBlock + threadIdx.x;
      b[u] = load_prim_box(prims, KIND, i < r.end ? i : r.end - 1);
    }
#pragma unroll
    for (int u = 0; u < kSegUnroll; u++) {
      const uint64_t wbase = t0 + uint64_t(u) * kSegBlock + uint64_t(wv) * kWave;
      if (wbase >= r.end)
        break; // uniform across the wave
      const uint64_t i = wbase + lane;
      const bool valid = i < r.end;
      float c[3] = { 0.0f, 0.0f, 0.0f };
      if (valid) {
        s = advance_scene(off, nseg, s, i);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const float sum = b[u].lo[k] + b[u].hi[k];
          c[k] = sum * 0.5f;
        }
        cx[i] = c[0];
        cy[i] = c[1];
        cz[i] = c[2];
      }
      const uint32_t s0 = __shfl(s, 0, kWave); // lane 0 is valid
      const bool same = !valid || s == s0;
      const bool uniform = ballot(!same) == 0ull;
      if (!(uniform && s0 == run)) {
        if (run != 0xffffffffu) {
          wave_reduce_min_max(rlo, rhi);
          if (lane == 0)
            flush_box(keys, run, rlo, rhi);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
          rlo[k] = INFINITY;
          rhi[k] = -INFINITY;
        }
        run = uniform ? s0 : 0xffffffffu;
        if (!uniform && valid)
          flush_box(keys, s, c, c); // a wave across a scene boundary
      }
      if (run != 0xffffffffu && valid) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          rlo[k] = std_min(rlo[k], c[k]);
          rhi[k] = std_max(rhi[k], c[k]);
        }
      }
    }
  }
  if (run != 0xffffffffu) {
    wave_reduce_min_max(rlo, rhi);
    if (lane == 0)
      flush_box(keys, run, rlo, rhi);
  }
}

__device__ __forceinline__ uint32_t
expand10_seg(uint32_t v)
{
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Morton codes with morton_domain<4> (D = 1024) against each scene's bounds,
// keyed (scene << 30 | code); LDS digit histograms of the radix passes
// first_pass .. passes - 1, one row per workgroup in part_hist.  counts !=
// nullptr: workgroup g encodes the chunk [g * chunk, (g + 1) * chunk) and
// also writes its pass-first_pass counts to counts[d * gridDim.x + g] (the
// reduce-then-scan sort's first chunk counts).  chunks != nullptr: workgroup
// g encodes [chunks[2g], chunks[2g + 1]) and writes counts[g * radix + d]
// (chunk-major, the scene-segmented pass), no part_hist row.
__global__ __launch_bounds__(kSegBlock) void
k_scene_encode(const float* __restrict__ cx, const float* __restrict__ cy,
               const float* __restrict__ cz, uint64_t n, const uint64_t* __restrict__ off,
               uint32_t nseg, const int32_t* __restrict__ bkeys, uint64_t* __restrict__ keys,
               uint32_t* __restrict__ part_hist, int digit_bits, int first_pass, int passes,
               uint32_t* __restrict__ counts, uint64_t chunk, const uint32_t* __restrict__ chunks)
{
  extern __shared__ uint32_t s_hist[];
  __shared__ uint32_t s_first;
  const int radix = 1 << digit_bits;
  const int words = radix * (passes - first_pass);
  for (int j = threadIdx.x; j < words; j += kSegBlock)
    s_hist[j] = 0u;
  BlockRange r;
  if (chunks) {
    r.begin = chunks[2 * blockIdx.x];
    r.end = chunks[2 * blockIdx.x + 1];
  } else if (counts) {
    r.begin = uint64_t(blockIdx.x) * chunk;
    if (r.begin > n)
      r.begin = n;
    r.end = (r.begin + chunk < n) ? r.begin + chunk : n;
  } else {
    r = block_range(n);
  }
  if (threadIdx.x == 0)
    s_first = r.begin < r.end ? find_scene(off, nseg, int64_t(r.begin), 1, false) : 0u;
  __syncthreads();
  uint32_t s = s_first;
  uint32_t cached = 0xffffffffu;
  float bmin[3] = { 0.0f, 0.0f, 0.0f }, scale[3] = { 0.0f, 0.0f, 0.0f };
  const uint64_t mask = uint64_t(radix - 1);
  for (uint64_t t0 = r.begin; t0 < r.end; t0 += uint64_t(kSegBlock) * kSegUnroll) {
    // unconditional loads (past the end: the range's last centre, unused)
    float x[kSegUnroll], y[kSegUnroll], z[kSegUnroll];
#pragma unroll
    for (int u = 0; u < kSegUnroll; u++) {
      const uint64_t i = t0 + uint64_t(u) * kSegBlock + threadIdx.x;
      const uint64_t j = i < r.end ? i : r.end - 1;
      x[u] = cx[j];
      y[u] = cy[j];
      z[u] = cz[j];
    }
#pragma unroll
    for (int u = 0; u < kSegUnroll; u++) {
      const uint64_t i = t0 + uint64_t(u) * kSegBlock + threadIdx.x;
      if (i >= r.end)
        continue;
      s = advance_scene(off, nseg, s, i);
      if (s != cached) {
        cached = s;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          bmin[k] = order_value(bkeys[6 * s + k]);
          const float ext = order_value(bkeys[6 * s + 3 + k]) - bmin[k];
          const float den = ext + 1.0f;
          scale[k] = 1024.0f / den;
        }
      }
      const float dx = x[u] - bmin[0];
      const float dy = y[u] - bmin[1];
      const float dz = z[u] - bmin[2];
      const float qx = dx * scale[0];
      const float qy = dy * scale[1];
      const float qz = dz * scale[2];
      const uint32_t code = (expand10_seg(uint32_t(qx)) << 2) | (expand10_seg(uint32_t(qy)) << 1) |
                            expand10_seg(uint32_t(qz));
      const uint64_t key = (uint64_t(s) << 30) | uint64_t(code);
      keys[i] = key;
      for (int p = first_pass; p < passes; p++)
        atomicAdd(&s_hist[(p - first_pass) * radix + int((key >> (p * digit_bits)) & mask)], 1u);
    }
  }
  __syncthreads();
  if (chunks) {
    for (int d = threadIdx.x; d < radix; d += kSegBlock)
      counts[size_t(blockIdx.x) * radix + d] = s_hist[d];
    return;
  }
  uint32_t* out = part_hist + size_t(blockIdx.x) * words;
  for (int j = threadIdx.x; j < words; j += kSegBlock)
    out[j] = s_hist[j];
  if (counts)
    for (int d = threadIdx.x; d < radix; d += kSegBlock)
      counts[size_t(d) * gridDim.x + blockIdx.x] = s_hist[d];
}

// hist[w] += sum over workgroups of part[g][w] (a slice of 16 per block row).
__global__ __launch_bounds__(kSegBlock) void
k_scene_hist_reduce(const uint32_t* __restrict__ part, uint32_t groups, uint32_t words,
                    uint32_t* __restrict__ hist)
{
  const uint32_t w = blockIdx.x * kSegBlock + threadIdx.x;
  if (w >= words)
    return;
  const uint32_t g0 = blockIdx.y * 16;
  uint32_t sum = 0;
  for (uint32_t g = g0; g < g0 + 16 && g < groups; g++)
    sum += part[size_t(g) * words + w];
  if (sum)
    atomicAdd(&hist[w], sum);
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

uint32_t
scene_encode_groups(uint64_t n)
{
  uint64_t g = (n + uint64_t(kSegBlock) * 16 - 1) / (uint64_t(kSegBlock) * 16);
  if (g < 1)
    g = 1;
  if (g > kSegMaxBlocks)
    g = kSegMaxBlocks;
  return uint32_t(g);
}

hipError_t
launch_scene_bounds_init(int32_t* bounds_key, uint32_t nseg, hipStream_t s)
{
  const uint32_t words = 6 * nseg;
  hipLaunchKernelGGL(k_scene_bounds_init, dim3((words + kSegBlock - 1) / kSegBlock),
                     dim3(kSegBlock), 0, s, bounds_key, nseg);
  return hipGetLastError();
}

hipError_t
launch_scene_bounds(const float* prims, int kind, uint64_t n, const uint64_t* prim_off,
                    uint32_t nseg, float* cx, float* cy, float* cz, int32_t* bounds_key,
                    hipStream_t s)
{
  const uint32_t g = scene_encode_groups(n);
  if (kind == kPrimTriangle)
    hipLaunchKernelGGL(k_scene_bounds<kPrimTriangle>, dim3(g), dim3(kSegBlock), 0, s, prims, n,
                       prim_off, nseg, cx, cy, cz, bounds_key);
  else
    hipLaunchKernelGGL(k_scene_bounds<kPrimAabb>, dim3(g), dim3(kSegBlock), 0, s, prims, n,
                       prim_off, nseg, cx, cy, cz, bounds_key);
  return hipGetLastError();
}

hipError_t
launch_scene_encode(const float* cx, const float* cy, const float* cz, uint64_t n,
                    const uint64_t* prim_off, uint32_t nseg, const int32_t* bounds_key,
                    uint64_t* keys, uint32_t* part_hist, uint32_t* hist, int digit_bits,
                    int first_pass, int passes, uint32_t* counts, uint32_t groups,
                    uint64_t chunk, const uint32_t* chunks, hipStream_t s)
{
  if (first_pass < 0 || first_pass >= passes)
    return hipErrorInvalidValue;
  const uint32_t words = uint32_t(1u << digit_bits) * uint32_t(passes - first_pass);
  if (words > uint32_t(kSegHistMaxWords))
    return hipErrorInvalidValue;
  uint32_t g = scene_encode_groups(n);
  if (chunks) {
    if (!counts || groups == 0 || passes != first_pass + 1)
      return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scene_encode, dim3(groups), dim3(kSegBlock), size_t(words) * 4, s, cx,
                       cy, cz, n, prim_off, nseg, bounds_key, keys, nullptr, digit_bits,
                       first_pass, passes, counts, 0, chunks);
    return hipGetLastError();
  }
  if (counts) {
    if (groups == 0 || chunk == 0 || uint64_t(groups) * chunk < n)
      return hipErrorInvalidValue;
    g = groups;
  }
  hipLaunchKernelGGL(k_scene_encode, dim3(g), dim3(kSegBlock), size_t(words) * 4, s, cx, cy, cz,
                     n, prim_off, nseg, bounds_key, keys, part_hist, digit_bits, first_pass,
                     passes, counts, chunk, nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_scene_hist_reduce, dim3((words + kSegBlock - 1) / kSegBlock, (g + 15) / 16),
                     dim3(kSegBlock), 0, s, part_hist, g, words,
                     hist + size_t(first_pass) * (size_t(1) << digit_bits));
  return hipGetLastError();
}

} // namespace lbvh_eng
