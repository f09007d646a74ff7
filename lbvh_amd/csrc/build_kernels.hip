// gfx950 build kernels: centroid bounds, Morton encode (+ the radix sort's
// digit histograms), Karras radix tree and the bottom-up box fit.  Each
// kernel cites the reference code whose results it reproduces bit for bit.
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

constexpr int kBlock = 256;
constexpr int kCentroidMaxBlocks = 1024;
constexpr int kEncodeMaxBlocks = 256;
// Largest subtree (in leaves) fitted inside one workgroup's LDS.  Nodes whose
// range holds more leaves form the "top" tree; their children with at most
// kTreelet leaves root the treelets.
constexpr int kTreelet = 1024;

// Parent link of a child (internal node c at parent[c], the leaf at sorted
// position k at parent[n - 1 + k]): the parent's index in bits 0-28 and
constexpr uint32_t kLinkGlobal = 1u << 31;  // the parent is a top node
constexpr uint32_t kLinkRight = 1u << 30;   // the child is the parent's right child
constexpr uint32_t kLinkSibLeaf = 1u << 29; // the child's sibling is a leaf
constexpr uint32_t kLinkIndex = kLinkSibLeaf - 1;

// ---- centroid bounds ---------------------------------------------------------
// Reproduces centroid_bounds_kernel + the serial union of per-thread boxes
// (lbvh.h:1282-1293, 1539-1543): the union of (min + max) * 0.5 over every
// primitive box (center_of, lbvh.h:1160-1167).  min/max are exact, so the
// reduction order cannot change the result for finite inputs.  Centres are
// also written out (SoA) for the encode pass.

__device__ __forceinline__ void
wave_reduce_box(float lo[3], float hi[3])
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float ol = __shfl_down(lo[k], off, kWave);
      float oh = __shfl_down(hi[k], off, kWave);
      lo[k] = std_min(lo[k], ol);
      hi[k] = std_max(hi[k], oh);
    }
  }
}

// Reduces one box per thread to thread 0 of the workgroup.
__device__ __forceinline__ void
block_reduce_box(float lo[3], float hi[3], float (*s_box)[6])
{
  wave_reduce_box(lo, hi);
  const int w = threadIdx.x / kWave;
  if (__lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_box[w][k] = lo[k];
      s_box[w][3 + k] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int ww = 1; ww < kBlock / kWave; ww++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = std_min(lo[k], s_box[ww][k]);
        hi[k] = std_max(hi[k], s_box[ww][3 + k]);
      }
  }
}

__global__ __launch_bounds__(kBlock) void
k_centroid_bounds(const float* __restrict__ prims, int kind, uint64_t n,
                  float* __restrict__ cx, float* __restrict__ cy,
                  float* __restrict__ cz, float* __restrict__ partials,
                  uint32_t* __restrict__ ticket, float* __restrict__ bounds,
                  uint32_t* __restrict__ zero_ptr, uint32_t zero_words)
{
  __shared__ float s_box[kBlock / kWave][6];
  __shared__ int s_last;
  const uint64_t tid = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;

  for (uint64_t j = tid; j < zero_words; j += stride)
    zero_ptr[j] = 0u;

  float lo[3] = { INFINITY, INFINITY, INFINITY };
  float hi[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (uint64_t i = tid; i < n; i += stride) {
    Box b = load_prim_box(prims, kind, i);
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float s = b.lo[k] + b.hi[k];
      c[k] = s * 0.5f;
      lo[k] = std_min(lo[k], c[k]);
      hi[k] = std_max(hi[k], c[k]);
    }
    cx[i] = c[0];
    cy[i] = c[1];
    cz[i] = c[2];
  }
  block_reduce_box(lo, hi, s_box);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      coherent_store_f32(&partials[6 * blockIdx.x + k], lo[k]);
      coherent_store_f32(&partials[6 * blockIdx.x + 3 + k], hi[k]);
    }
    vmem_drain();
    uint32_t t = atomic_add_agent(ticket, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last)
    return;
  // Last workgroup: union of the per-workgroup boxes.
  float glo[3] = { INFINITY, INFINITY, INFINITY };
  float ghi[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (uint32_t g = threadIdx.x; g < gridDim.x; g += kBlock)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      glo[k] = std_min(glo[k], coherent_load_f32(&partials[6 * g + k]));
      ghi[k] = std_max(ghi[k], coherent_load_f32(&partials[6 * g + 3 + k]));
    }
  __syncthreads(); // s_box reuse
  block_reduce_box(glo, ghi, s_box);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      bounds[k] = glo[k];
      bounds[3 + k] = ghi[k];
    }
    *ticket = 0u;
  }
}

// ---- Morton encode --------------------------------------------------------------
// morton_curve_kernel (lbvh.h:1432-1473): scale = D / (extent + 1) with
// D = 1024 (morton_domain<4>) or 2^20 (morton_domain<8>); q = trunc((c - min)
// * scale); code = expand(qx) << 2 | expand(qy) << 1 | expand(qz) with the
// magic-mask expansions of morton_encoder<4>/<8> (lbvh.h:1352-1358,
// 1377-1385).  The 63-bit variant applies morton_encoder<8> to FP32 centres.
// Each workgroup also counts the digits of every radix pass in LDS and writes
// its counts (plain stores) for k_hist_reduce.

__device__ __forceinline__ uint32_t
expand10(uint32_t n)
{
  n = (n | (n << 16)) & 0x030000ffu;
  n = (n | (n << 8)) & 0x0300f00fu;
  n = (n | (n << 4)) & 0x030c30c3u;
  n = (n | (n << 2)) & 0x09249249u;
  return n;
}

__device__ __forceinline__ uint64_t
expand20(uint64_t n)
{
  n = (n | n << 32) & 0x001f00000000ffffull;
  n = (n | n << 16) & 0x001f0000ff0000ffull;
  n = (n | n << 8) & 0x100f00f00f00f00full;
  n = (n | n << 4) & 0x10c30c30c30c30c3ull;
  n = (n | n << 2) & 0x1249249249249249ull;
  return n;
}

template<typename KeyT>
__device__ __forceinline__ KeyT
encode(float qx, float qy, float qz);

template<>
__device__ __forceinline__ uint32_t
encode<uint32_t>(float qx, float qy, float qz)
{
  uint32_t x = uint32_t(qx), y = uint32_t(qy), z = uint32_t(qz);
  return (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
}

template<>
__device__ __forceinline__ uint64_t
encode<uint64_t>(float qx, float qy, float qz)
{
  uint64_t x = uint64_t(qx), y = uint64_t(qy), z = uint64_t(qz);
  return (expand20(x) << 2) | (expand20(y) << 1) | expand20(z);
}

constexpr int kHistMaxWords = 8192; // passes * radix, 32 KiB of LDS
constexpr int kEncodeUnroll = 4;

template<typename KeyT>
__global__ __launch_bounds__(kBlock) void
k_encode(const float* __restrict__ cx, const float* __restrict__ cy,
         const float* __restrict__ cz, uint64_t n,
         const float* __restrict__ bounds, KeyT* __restrict__ codes,
         uint32_t* __restrict__ part_hist, int digit_bits, int passes)
{
  extern __shared__ uint32_t s_hist[];
  const int radix = 1 << digit_bits;
  const int words = radix * passes;
  for (int j = threadIdx.x; j < words; j += kBlock)
    s_hist[j] = 0u;
  const float dom = sizeof(KeyT) == 4 ? 1024.0f : 1048576.0f;
  float bmin[3], scale[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bmin[k] = bounds[k];
    float ext = bounds[3 + k] - bmin[k];
    float den = ext + 1.0f;
    scale[k] = dom / den;
  }
  __syncthreads();
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  const KeyT mask = KeyT(radix - 1);
  for (uint64_t i0 = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i0 < n;
       i0 += stride * kEncodeUnroll) {
    float x[kEncodeUnroll], y[kEncodeUnroll], z[kEncodeUnroll];
#pragma unroll
    for (int u = 0; u < kEncodeUnroll; u++) {
      uint64_t i = i0 + u * stride;
      if (i < n) {
        x[u] = cx[i];
        y[u] = cy[i];
        z[u] = cz[i];
      }
    }
#pragma unroll
    for (int u = 0; u < kEncodeUnroll; u++) {
      uint64_t i = i0 + u * stride;
      if (i >= n)
        continue;
      float dx = x[u] - bmin[0];
      float dy = y[u] - bmin[1];
      float dz = z[u] - bmin[2];
      float qx = dx * scale[0];
      float qy = dy * scale[1];
      float qz = dz * scale[2];
      KeyT code = encode<KeyT>(qx, qy, qz);
      codes[i] = code;
      for (int p = 0; p < passes; p++) {
        uint32_t d = uint32_t((code >> (p * digit_bits)) & mask);
        atomicAdd(&s_hist[p * radix + d], 1u);
      }
    }
  }
  __syncthreads();
  uint32_t* out = part_hist + size_t(blockIdx.x) * words;
  for (int j = threadIdx.x; j < words; j += kBlock)
    out[j] = s_hist[j];
}

// Sums the per-workgroup digit counts into the (zeroed) histogram:
// hist[w] += sum of part[g][w] over a slice of kHistSlice workgroups.
constexpr uint32_t kHistSlice = 16;

__global__ __launch_bounds__(kBlock) void
k_hist_reduce(const uint32_t* __restrict__ part, uint32_t groups, uint32_t words,
              uint32_t* __restrict__ hist)
{
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= words)
    return;
  const uint32_t g0 = blockIdx.y * kHistSlice;
  uint32_t s = 0;
#pragma unroll
  for (uint32_t g = 0; g < kHistSlice; g++)
    if (g0 + g < groups)
      s += part[size_t(g0 + g) * words + w];
  if (s)
    atomicAdd(&hist[w], s);
}

// ---- Karras radix tree -----------------------------------------------------------
// divide_node + builder_kernel (lbvh.h:1588-1688, 1713-1731), one internal
// node per thread:
//   delta(j, k) = -1 if k is outside [0, n); clz(code_j ^ code_k) if the codes
//   differ; clz(j ^ k) + bits(code) if they are equal;
//   d = sign(delta(i, i+1) - delta(i, i-1)); delta_min = delta(i, i-d);
//   l_max starts at 128 and grows x4 while delta(i, i + l_max d) > delta_min
//   (the loop exits on k >= n only); binary search for l; split search with
//   ceil_div; gamma = i + s d + min(d, 0).
// left = gamma | leaf if min(i, j) == gamma; right = gamma + 1 | leaf if
// max(i, j) == gamma + 1.  With PRIMITIVE_ID leaves, a leaf stores the id of
// the primitive at that sorted position.
//
// The codes a workgroup's nodes search within +-128 positions are staged in
// LDS; farther probes (large ranges near the root) read global memory.
// Besides left/right it writes, for the fit: parent links (see kLink*); the
// primitive of each leaf position (primitive-id leaves); and clears the fit
// flags.

template<typename CodeT>
struct CodeTraits;
template<>
struct CodeTraits<uint32_t>
{
  static constexpr int bits = 32;
  __device__ static int clz(uint32_t x) { return __clz(int(x)); }
};
template<>
struct CodeTraits<uint64_t>
{
  static constexpr int bits = 64;
  __device__ static int clz(uint64_t x) { return __clzll((long long)x); }
};

// Sorted keys hold the Morton code alone (one scene) or (scene << 30 | code)
// for a batch of scenes with 30-bit codes; the tree compares codes only.
template<typename KeyT, typename CodeT, bool SEG>
__device__ __forceinline__ CodeT
key_code(KeyT k)
{
  if constexpr (SEG)
    return CodeT(k & 0x3fffffffull);
  else
    return CodeT(k);
}

constexpr int kKarrasHalo = 128;
constexpr int kKarrasWindow = kBlock + 2 * kKarrasHalo + 64; // + room for scene boundaries

// Codes of one scene (positions base .. base + n - 1 of the sorted array),
// served from an LDS copy of sorted positions [w0, w0 + kKarrasWindow) when
// possible.  Positions passed in are scene-local.
template<typename KeyT, typename CodeT, bool SEG>
struct CodeWindow
{
  const KeyT* __restrict__ g;
  const KeyT* s;
  int64_t w0;
  int64_t base;
  int64_t n;
  __device__ CodeT at(int64_t k) const
  {
    const int64_t x = base + k;
    const int64_t o = x - w0;
    return key_code<KeyT, CodeT, SEG>((o >= 0 && o < kKarrasWindow) ? s[o] : g[x]);
  }
  // calc_delta (lbvh.h:1602-1621) over scene-local positions
  __device__ int delta(int64_t j, CodeT cj, int64_t k) const
  {
    if (k < 0 || k >= n)
      return -1;
    CodeT ck = at(k);
    if (cj == ck)
      return CodeTraits<CodeT>::clz(CodeT(uint64_t(j ^ k))) + CodeTraits<CodeT>::bits;
    return CodeTraits<CodeT>::clz(cj ^ ck);
  }
};

// Largest x in [0, span) with pred(x), for a predicate that holds on a
// prefix of [1, span) and fails after it (x = 0 when pred(1) fails); span is
// a power of two.  8-way steps issue 7 independent probes each.
template<class Pred>
__device__ __forceinline__ int64_t
search_prefix(int64_t span, const Pred& pred)
{
  int64_t x = 0;
  while (span >= 8) {
    const int64_t step = span >> 3;
    int cnt = 0;
#pragma unroll
    for (int q = 1; q < 8; q++)
      cnt += pred(x + q * step) ? 1 : 0;
    x += cnt * step;
    span = step;
  }
  while (span > 1) {
    const int64_t t = span >> 1;
    if (pred(x + t))
      x += t;
    span = t;
  }
  return x;
}

// The same result by halving steps.
template<class Pred>
__device__ __forceinline__ int64_t
search_binary(int64_t span, const Pred& pred)
{
  int64_t x = 0;
  while (span > 1) {
    const int64_t t = span >> 1;
    if (pred(x + t))
      x += t;
    span = t;
  }
  return x;
}

template<typename KeyT, typename CodeT, bool SEG>
__global__ __launch_bounds__(kBlock) void
k_karras(const KeyT* __restrict__ keys, const uint32_t* __restrict__ ids, Scenes sc,
         Node* __restrict__ nodes, uint32_t* __restrict__ parent,
         uint32_t* __restrict__ leafprim, uint32_t* __restrict__ flags,
         uint32_t* __restrict__ other_end, uint2* __restrict__ treelets,
         uint32_t* __restrict__ treelet_count, WideNode* __restrict__ wide)
{
  __shared__ KeyT s_keys[kKarrasWindow];
  // batches: offsets of the scenes this workgroup's nodes can belong to (a
  // scene has at least one node, so at most kBlock scenes, + 1 end offset)
  __shared__ uint64_t s_off[SEG ? kBlock + 1 : 1];
  __shared__ uint32_t s_first;
  const int64_t nodes_total = sc.n - (SEG ? int64_t(sc.nseg) : 1);
  const int64_t g0 = int64_t(blockIdx.x) * kBlock;
  const int64_t g = g0 + threadIdx.x;
  uint32_t s0 = 0;
  uint32_t span = 0; // scenes staged in s_off (entries span + 1)
  if constexpr (SEG) {
    if (threadIdx.x == 0)
      s_first = find_scene(sc.prim_off, sc.nseg, g0, 1, true);
    __syncthreads();
    s0 = s_first;
    span = sc.nseg - s0 < uint32_t(kBlock) ? sc.nseg - s0 : uint32_t(kBlock);
    for (uint32_t t = threadIdx.x; t <= span; t += kBlock)
      s_off[t] = sc.prim_off[s0 + t];
  }
  // scene of the workgroup's first node sets the window (sorted position of
  // node g of scene s is g + s)
  const int64_t w0 = g0 + s0 - kKarrasHalo;
  for (int t = threadIdx.x; t < kKarrasWindow; t += kBlock) {
    const int64_t x = w0 + t;
    s_keys[t] = (x >= 0 && x < sc.n) ? keys[x] : KeyT(0);
  }
  __syncthreads();
  if (g >= nodes_total)
    return;
  uint32_t seg = 0;
  int64_t P = 0, n = sc.n;
  if constexpr (SEG) {
    // largest j < span with node base s_off[j] - (s0 + j) <= g
    uint32_t lo = 0, hi = span;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (int64_t(s_off[mid]) - int64_t(s0 + mid) <= g)
        lo = mid;
      else
        hi = mid;
    }
    seg = s0 + lo;
    P = int64_t(s_off[lo]);
    n = int64_t(s_off[lo + 1]) - P;
  }
  const int64_t node_base = P - seg;
  const int64_t link_base = 2 * P - seg;
  const int64_t i = g - node_base;
  if (n < 2 || i < 0 || i >= n - 1)
    return; // scenes have at least 2 primitives (checked by the host)
  const CodeWindow<KeyT, CodeT, SEG> cw{ keys, s_keys, w0, P, n };
  const CodeT ci = cw.at(i);
  const int dd = cw.delta(i, ci, i + 1) - cw.delta(i, ci, i - 1);
  const int64_t d = (dd < 0) ? -1 : 1;
  const int delta_min = cw.delta(i, ci, i - d);
  // The reference grows l_max from 128 by x4 while delta(i, i + l_max d)
  // exceeds delta_min (stopping when i + l_max d >= n), then binary-searches
  // the largest l < l_max with delta(i, i + l d) > delta_min, then the
  // largest split s < l with delta(i, i + s d) > delta_node.  For sorted
  // keys delta(i, k) does not grow as k moves away from i, so each of these
  // predicates is true on a prefix and false after it: the searches below
  // find the same boundaries, with independent probes (all l_max candidates
  // at once, then 7 probes per 8-way step) where the span leaves the LDS
  // window.
  auto candidate = [&](int q) {
    const int64_t span = int64_t(128) << (2 * q);
    const int64_t k = i + (d > 0 ? span : -span);
    return (k < n) && (cw.delta(i, ci, k) > delta_min);
  };
  int m = 0;
  if (candidate(0)) { // most nodes stop here, inside the LDS window
    m = 1;
#pragma unroll
    for (int q = 1; q < 13; q++)
      m += candidate(q) ? 1 : 0; // prefix-true: m = index of the first failing candidate
  }
  const int64_t l_max = int64_t(128) << (2 * m);
  auto pred_l = [&](int64_t x) { return cw.delta(i, ci, i + (d > 0 ? x : -x)) > delta_min; };
  const int64_t l = (m == 0) ? search_binary(l_max, pred_l) : search_prefix(l_max, pred_l);
  const int64_t j = i + (d > 0 ? l : -l);
  const int delta_node = cw.delta(i, ci, j);
  int64_t span_s = 1;
  while (span_s < l)
    span_s <<= 1;
  auto pred_s = [&](int64_t x) { return cw.delta(i, ci, i + (d > 0 ? x : -x)) > delta_node; };
  const int64_t s = (span_s <= kKarrasHalo) ? search_binary(span_s, pred_s)
                                            : search_prefix(span_s, pred_s);
  const int64_t gamma = (d > 0) ? (i + s) : (i - s - 1);
  if (gamma < 0 || gamma + 1 > n - 1)
    return; // unreachable for sorted codes; keeps every index below in range
  const int64_t mn = (i < j) ? i : j;
  const int64_t mx = (j > i) ? j : i;
  const bool lleaf = (mn == gamma);
  const bool rleaf = (mx == gamma + 1);
  // a top node: its children hand their boxes up through the top-tree pass
  const bool top = (mx - mn + 1) > kTreelet;
  const uint32_t up = uint32_t(i) | (top ? kLinkGlobal : 0u);
  const uint32_t left_link = up | (rleaf ? kLinkSibLeaf : 0u);
  const uint32_t right_link = up | kLinkRight | (lleaf ? kLinkSibLeaf : 0u);
  other_end[node_base + i] = uint32_t(j);
  if (top) {
    // children with at most kTreelet leaves root treelets; an entry holds the
    // leaf range as sorted positions of the whole batch, and in bit 31 of .y
    // whether its root is a right child (the root is numbered hi for a left
    // child, lo for a right child)
    if (gamma - mn + 1 <= kTreelet)
      treelets[atomicAdd(treelet_count, 1u)] = make_uint2(uint32_t(P + mn), uint32_t(P + gamma));
    if (mx - gamma <= kTreelet)
      treelets[atomicAdd(treelet_count, 1u)] =
        make_uint2(uint32_t(P + gamma + 1), uint32_t(P + mx) | 0x80000000u);
  } else if (i == 0) {
    treelets[atomicAdd(treelet_count, 1u)] = make_uint2(uint32_t(P), uint32_t(P + n - 1));
  }
  // left / right are written by the fit together with the box (one full
  // node and one full traversal record per internal node)
  if (lleaf) {
    parent[link_base + (n - 1) + gamma] = left_link;
    if (leafprim)
      leafprim[P + gamma] = uint32_t(int64_t(ids[P + gamma]) - P);
  } else {
    parent[link_base + gamma] = left_link;
  }
  if (rleaf) {
    parent[link_base + (n - 1) + gamma + 1] = right_link;
    if (leafprim)
      leafprim[P + gamma + 1] = uint32_t(int64_t(ids[P + gamma + 1]) - P);
  } else {
    parent[link_base + gamma + 1] = right_link;
  }
  if (top)
    flags[node_base + i] = 0u; // only top nodes use device-scope visit flags
}

// ---- bottom-up fit -------------------------------------------------------------
// builder::fit_boxes (lbvh.h:1827-1866) computes every internal box as
// union(left box, right box), left operand first, where a leaf's box is the
// converter applied to the primitive its index names.  Here the second of
// the two child computations to finish computes the parent's box.  In the
// Karras layout the children of a node with split g are g and g + 1 (leaf
// positions or node indices), so the parent link's side and sibling-is-leaf
// bits locate the sibling without reading the node.
//
// Pass 1 (k_fit_treelets): each treelet -- a subtree of at most kTreelet
// leaves under a top node -- is fitted by one workgroup entirely in LDS: the
// leaves' boxes and the internal nodes' links are staged once, then one
// thread per leaf climbs with LDS visit flags.  Workgroups pull treelets
// from the list the tree kernel wrote.
// Pass 2 (k_fit_top): one thread per treelet climbs from the treelet root
// through the top nodes with device-scope visit flags and device-coherent
// box stores/loads (the top tree is shallow).  Flags are cleared by their
// second arrival, so a refit needs no memset.

__device__ __forceinline__ Box
load_node_box_coherent(const Node* nd)
{
  Box b;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    b.lo[k] = coherent_load_f32(&nd->bmin[k]);
    b.hi[k] = coherent_load_f32(&nd->bmax[k]);
  }
  return b;
}

__device__ __forceinline__ void
store_node_box_coherent(Node* nd, const Box& b)
{
#pragma unroll
  for (int k = 0; k < 3; k++) {
    coherent_store_f32(&nd->bmin[k], b.lo[k]);
    coherent_store_f32(&nd->bmax[k], b.hi[k]);
  }
}

// A completed internal node: box and children, one full 32-byte record.
__device__ __forceinline__ void
store_node(Node* nd, const Box& b, uint32_t left, uint32_t right)
{
  uint4* q = reinterpret_cast<uint4*>(nd);
  q[0] = make_uint4(__float_as_uint(b.lo[0]), __float_as_uint(b.lo[1]), __float_as_uint(b.lo[2]),
                    __float_as_uint(b.hi[0]));
  q[1] = make_uint4(__float_as_uint(b.hi[1]), __float_as_uint(b.hi[2]), left, right);
}

// Its traversal record: both child boxes and both children, one full
// 64-byte record.
__device__ __forceinline__ void
store_wide(WideNode* w, const Box& lb, const Box& rb, uint32_t left, uint32_t right)
{
  uint4* q = reinterpret_cast<uint4*>(w);
  q[0] = make_uint4(__float_as_uint(lb.lo[0]), __float_as_uint(lb.lo[1]),
                    __float_as_uint(lb.lo[2]), __float_as_uint(lb.hi[0]));
  q[1] = make_uint4(__float_as_uint(lb.hi[1]), __float_as_uint(lb.hi[2]),
                    __float_as_uint(rb.lo[0]), __float_as_uint(rb.lo[1]));
  q[2] = make_uint4(__float_as_uint(rb.lo[2]), __float_as_uint(rb.hi[0]),
                    __float_as_uint(rb.hi[1]), __float_as_uint(rb.hi[2]));
  q[3] = make_uint4(left, right, 0u, 0u);
}

// Child value of a node: an internal index, or a leaf (bit 31) holding the
// sorted position (reference leaves) or the primitive id there.
__device__ __forceinline__ uint32_t
child_value(int64_t idx, bool is_leaf, const uint32_t* __restrict__ leafprim)
{
  if (!is_leaf)
    return uint32_t(idx);
  return (leafprim ? leafprim[idx] : uint32_t(idx)) | kLeafBit;
}

__device__ __forceinline__ Box
leaf_box(const float* __restrict__ prims, int kind, const uint32_t* __restrict__ leafprim,
         int64_t pos)
{
  const uint32_t v = leafprim ? leafprim[pos] : uint32_t(pos);
  return load_prim_box(prims, kind, uint64_t(v));
}

// The scene of a treelet entry (batch positions) and its arrays.
struct TreeletScene
{
  int64_t lo, hi, n; // scene-local leaf range, scene size
  int64_t P, nb, lb;
};

__device__ __forceinline__ TreeletScene
treelet_scene(const Scenes& sc, uint2 e)
{
  const int64_t glo = e.x;
  const int64_t ghi = e.y & 0x7fffffffu;
  const uint32_t seg = sc.prim_off ? find_scene(sc.prim_off, sc.nseg, glo, 1, false) : 0u;
  const SceneView v = scene_view(sc, seg);
  return TreeletScene{ glo - v.P, ghi - v.P, v.n, v.P, v.nb, v.lb };
}

__global__ __launch_bounds__(kTreelet) void
k_fit_treelets(const float* __restrict__ prims_all, int kind, Scenes sc, Node* __restrict__ nodes_all,
               const uint32_t* __restrict__ parent_all, const uint32_t* __restrict__ leafprim_all,
               const uint2* __restrict__ treelets, const uint32_t* __restrict__ count_ptr,
               uint32_t* __restrict__ queue, WideNode* __restrict__ wide_all)
{
  __shared__ uint32_t s_flag[kTreelet];
  __shared__ uint32_t s_link[kTreelet]; // link of internal node lo + t
  __shared__ Box s_leaf[kTreelet];      // box of the leaf at position lo + t
  __shared__ Box s_box[kTreelet];       // box of internal node lo + t
  __shared__ uint32_t s_leafv[kTreelet]; // child value of the leaf at position lo + t
  __shared__ uint32_t s_item;
  const int t = threadIdx.x;
  const uint32_t count = *count_ptr;
  for (;;) {
    if (t == 0)
      s_item = atomic_add_agent(queue, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= count)
      return; // uniform across the workgroup
    const TreeletScene ts = treelet_scene(sc, treelets[item]);
    const int64_t lo = ts.lo, hi = ts.hi, n = ts.n;
    const int64_t m = hi - lo + 1;
    if (lo >= 0 && m >= 2 && m <= kTreelet && hi < n) {
      const float* prims = prims_all + prim_stride(kind) * ts.P;
      Node* nodes = nodes_all + ts.nb;
      const uint32_t* parent = parent_all + ts.lb;
      const uint32_t* leafprim = leafprim_all ? leafprim_all + ts.P : nullptr;
      WideNode* wide = wide_all + ts.nb;
      const int64_t k = lo + t;
      uint32_t link = 0u;
      if (t < m) {
        s_leaf[t] = leaf_box(prims, kind, leafprim, k);
        s_leafv[t] = child_value(k, true, leafprim);
        link = parent[(n - 1) + k];
        if (k < n - 1)
          s_link[t] = parent[k];
      }
      s_flag[t] = 0u;
      __syncthreads();
      if (t < m) {
        Box mine = s_leaf[t];
        int64_t cur = k;
        uint32_t cur_v = s_leafv[t]; // child value of cur
        for (;;) {
          if (link & kLinkGlobal)
            break; // cur is the treelet root
          const int64_t p = int64_t(link & kLinkIndex);
          const bool right = (link & kLinkRight) != 0;
          const int64_t sib = right ? cur - 1 : cur + 1;
          // links come from the tree kernel; never taken for a valid tree
          if (p < lo || p > hi || p >= n - 1 || sib < lo || sib > hi)
            break;
          const uint32_t old = __hip_atomic_fetch_add(&s_flag[p - lo], 1u, __ATOMIC_ACQ_REL,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP);
          if (old == 0u)
            break; // first arrival: the sibling's thread completes p
          const bool sib_leaf = (link & kLinkSibLeaf) != 0;
          const Box ob = sib_leaf ? s_leaf[sib - lo] : s_box[sib - lo];
          const uint32_t sib_v = sib_leaf ? s_leafv[sib - lo] : uint32_t(sib);
          const Box& lb = right ? ob : mine;
          const Box& rb = right ? mine : ob;
          const uint32_t lv = right ? sib_v : cur_v;
          const uint32_t rv = right ? cur_v : sib_v;
          const Box b = box_union(lb, rb);
          store_node(&nodes[p], b, lv, rv);
          store_wide(&wide[p], lb, rb, lv, rv);
          s_box[p - lo] = b;
          if (p == 0)
            break; // the whole tree was one treelet
          mine = b;
          cur = p;
          cur_v = uint32_t(p);
          link = s_link[p - lo];
        }
      }
    }
    __syncthreads(); // LDS and s_item are reused by the next treelet
  }
}

__global__ __launch_bounds__(kBlock) void
k_fit_top(const float* __restrict__ prims_all, int kind, Scenes sc, Node* __restrict__ nodes_all,
          const uint32_t* __restrict__ parent_all, const uint32_t* __restrict__ leafprim_all,
          const uint2* __restrict__ treelets, const uint32_t* __restrict__ count_ptr,
          uint32_t* __restrict__ flags_all, WideNode* __restrict__ wide_all)
{
  const uint32_t count = *count_ptr;
  for (uint32_t item = blockIdx.x * kBlock + threadIdx.x; item < count;
       item += gridDim.x * kBlock) {
    const uint2 e = treelets[item];
    const TreeletScene ts = treelet_scene(sc, e);
    const int64_t lo = ts.lo, hi = ts.hi, n = ts.n;
    if (lo < 0 || hi >= n || lo > hi || (lo == 0 && hi == n - 1))
      continue; // the whole tree was one treelet (or a corrupt entry)
    const float* prims = prims_all + prim_stride(kind) * ts.P;
    Node* nodes = nodes_all + ts.nb;
    const uint32_t* parent = parent_all + ts.lb;
    const uint32_t* leafprim = leafprim_all ? leafprim_all + ts.P : nullptr;
    uint32_t* flags = flags_all + ts.nb;
    WideNode* wide = wide_all + ts.nb;
    Box mine;
    int64_t cur;
    uint32_t cur_v;
    uint32_t link;
    if (lo == hi) {
      cur = lo;
      cur_v = child_value(lo, true, leafprim);
      mine = leaf_box(prims, kind, leafprim, lo);
      link = parent[(n - 1) + lo];
    } else {
      cur = (e.y & 0x80000000u) ? lo : hi;
      if (cur >= n - 1)
        continue;
      cur_v = uint32_t(cur);
      const float4 a = *reinterpret_cast<const float4*>(&nodes[cur].bmin[0]);
      const float2 c = *reinterpret_cast<const float2*>(&nodes[cur].bmax[1]);
      mine.lo[0] = a.x;
      mine.lo[1] = a.y;
      mine.lo[2] = a.z;
      mine.hi[0] = a.w;
      mine.hi[1] = c.x;
      mine.hi[2] = c.y;
      link = parent[cur];
    }
    for (;;) {
      const int64_t p = int64_t(link & kLinkIndex);
      const bool right = (link & kLinkRight) != 0;
      const int64_t sib = right ? cur - 1 : cur + 1;
      if (!(link & kLinkGlobal) || p >= n - 1 || sib < 0 || sib >= n)
        break; // never taken for a valid tree
      if (atomic_add_agent(&flags[p], 1u) == 0u)
        break; // first arrival: the sibling's thread completes p
      const bool sib_leaf = (link & kLinkSibLeaf) != 0;
      Box ob;
      if (sib_leaf) {
        ob = leaf_box(prims, kind, leafprim, sib);
      } else {
        if (sib >= n - 1)
          break;
        ob = load_node_box_coherent(&nodes[sib]);
      }
      const uint32_t sib_v = child_value(sib, sib_leaf, leafprim);
      const Box& lb = right ? ob : mine;
      const Box& rb = right ? mine : ob;
      const uint32_t lv = right ? sib_v : cur_v;
      const uint32_t rv = right ? cur_v : sib_v;
      const Box b = box_union(lb, rb);
      coherent_store_u32(&flags[p], 0u);
      store_wide(&wide[p], lb, rb, lv, rv);
      if (p == 0) {
        store_node(&nodes[0], b, lv, rv);
        break;
      }
      // the box is read by the thread completing p's parent, possibly on
      // another XCD
      store_node_box_coherent(&nodes[p], b);
      nodes[p].left = lv;
      nodes[p].right = rv;
      vmem_drain();
      mine = b;
      cur = p;
      cur_v = uint32_t(p);
      link = parent[p];
    }
  }
}

// Rebuilds the treelet list of an existing tree (refit): every child that is
// not a top node but whose parent is roots a treelet; its leaf range is
// [c, other_end[c]] (either order) for an internal node, [k, k] for a leaf.
// One thread per parent link of the batch.
__global__ __launch_bounds__(kBlock) void
k_enum_treelets(Scenes sc, const uint32_t* __restrict__ parent_all,
                const uint32_t* __restrict__ other_end_all, uint2* __restrict__ treelets,
                uint32_t* __restrict__ count)
{
  const int64_t links = 2 * sc.n - (sc.prim_off ? int64_t(sc.nseg) : 1);
  const int64_t gx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (gx >= links)
    return;
  const uint32_t seg = sc.prim_off ? find_scene(sc.prim_off, sc.nseg, gx, 2, true) : 0u;
  const SceneView v = scene_view(sc, seg);
  const int64_t n = v.n;
  const int64_t x = gx - v.lb;
  if (x < 0 || x >= 2 * n - 1)
    return;
  const uint32_t P = uint32_t(v.P);
  if (x == 0) {
    if (n <= kTreelet)
      treelets[atomicAdd(count, 1u)] = make_uint2(P, P + uint32_t(n - 1));
    return;
  }
  const uint32_t link = parent_all[v.lb + x];
  if (!(link & kLinkGlobal))
    return;
  if (x < n - 1) {
    const int64_t j = other_end_all[v.nb + x];
    const int64_t lo = x < j ? x : j;
    const int64_t hi = x < j ? j : x;
    if (hi - lo + 1 > kTreelet)
      return; // a top node itself
    treelets[atomicAdd(count, 1u)] = make_uint2(
      P + uint32_t(lo), (P + uint32_t(hi)) | ((link & kLinkRight) ? 0x80000000u : 0u));
  } else {
    const uint32_t k = P + uint32_t(x - (n - 1));
    treelets[atomicAdd(count, 1u)] = make_uint2(k, k);
  }
}

__global__ void
k_fill_u32(uint32_t* __restrict__ p, uint64_t words, uint32_t value)
{
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride)
    p[i] = value;
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

static inline uint32_t
grid_for(uint64_t work, uint64_t per_block, uint32_t cap)
{
  uint64_t g = (work + per_block - 1) / per_block;
  if (g < 1)
    g = 1;
  if (g > cap)
    g = cap;
  return uint32_t(g);
}

uint32_t
centroid_partials_needed(uint64_t n)
{
  return grid_for(n, uint64_t(kBlock) * 4, kCentroidMaxBlocks);
}

uint32_t
encode_groups(uint64_t n)
{
  return grid_for(n, uint64_t(kBlock) * 16, kEncodeMaxBlocks);
}

hipError_t
launch_centroid_bounds(const float* prims, int kind, uint64_t n, float* cx, float* cy,
                       float* cz, float* partials, uint32_t* ticket, float* bounds,
                       uint32_t* zero_ptr, uint32_t zero_words, hipStream_t s)
{
  uint32_t g = centroid_partials_needed(n);
  hipLaunchKernelGGL(k_centroid_bounds, dim3(g), dim3(kBlock), 0, s, prims, kind, n, cx, cy,
                     cz, partials, ticket, bounds, zero_ptr, zero_words);
  return hipGetLastError();
}

hipError_t
launch_encode(int code_bits, const float* cx, const float* cy, const float* cz, uint64_t n,
              const float* bounds, void* codes, uint32_t* part_hist, uint32_t* hist,
              int digit_bits, int passes, hipStream_t s)
{
  const uint32_t words = uint32_t(1u << digit_bits) * uint32_t(passes);
  if (words > uint32_t(kHistMaxWords))
    return hipErrorInvalidValue;
  size_t lds = size_t(words) * sizeof(uint32_t);
  uint32_t g = encode_groups(n);
  if (code_bits == 30)
    hipLaunchKernelGGL(k_encode<uint32_t>, dim3(g), dim3(kBlock), lds, s, cx, cy, cz, n,
                       bounds, static_cast<uint32_t*>(codes), part_hist, digit_bits, passes);
  else
    hipLaunchKernelGGL(k_encode<uint64_t>, dim3(g), dim3(kBlock), lds, s, cx, cy, cz, n,
                       bounds, static_cast<uint64_t*>(codes), part_hist, digit_bits, passes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_hist_reduce,
                     dim3((words + kBlock - 1) / kBlock, (g + kHistSlice - 1) / kHistSlice),
                     dim3(kBlock), 0, s, part_hist, g, words, hist);
  return hipGetLastError();
}

static inline Scenes
make_scenes(const uint64_t* prim_off, uint32_t nseg, uint64_t n)
{
  Scenes sc;
  sc.prim_off = prim_off;
  sc.nseg = prim_off ? nseg : 1u;
  sc.n = int64_t(n);
  return sc;
}

hipError_t
launch_karras(int code_bits, const void* sorted, const uint32_t* ids, uint64_t n,
              const uint64_t* prim_off, uint32_t nseg, void* nodes, uint32_t* parent,
              uint32_t* leafprim, uint32_t* flags, uint32_t* other_end, void* treelets,
              uint32_t* treelet_count, void* wide, hipStream_t s)
{
  const Scenes sc = make_scenes(prim_off, nseg, n);
  const uint64_t m = n - sc.nseg;
  const uint32_t g = uint32_t((m + kBlock - 1) / kBlock);
  uint2* tl = static_cast<uint2*>(treelets);
  Node* nd = static_cast<Node*>(nodes);
  WideNode* wd = static_cast<WideNode*>(wide);
  if (prim_off) {
    if (code_bits != 30)
      return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_karras<uint64_t, uint32_t, true>), dim3(g), dim3(kBlock), 0, s,
                       static_cast<const uint64_t*>(sorted), ids, sc, nd, parent, leafprim, flags,
                       other_end, tl, treelet_count, wd);
  } else if (code_bits == 30) {
    hipLaunchKernelGGL((k_karras<uint32_t, uint32_t, false>), dim3(g), dim3(kBlock), 0, s,
                       static_cast<const uint32_t*>(sorted), ids, sc, nd, parent, leafprim, flags,
                       other_end, tl, treelet_count, wd);
  } else {
    hipLaunchKernelGGL((k_karras<uint64_t, uint64_t, false>), dim3(g), dim3(kBlock), 0, s,
                       static_cast<const uint64_t*>(sorted), ids, sc, nd, parent, leafprim, flags,
                       other_end, tl, treelet_count, wd);
  }
  return hipGetLastError();
}

hipError_t
launch_fit(const float* prims, int kind, uint64_t n, const uint64_t* prim_off, uint32_t nseg,
           void* nodes, const uint32_t* parent, const uint32_t* leafprim, uint32_t* flags,
           const void* treelets, uint32_t* treelet_count, uint32_t* queue, void* wide,
           hipStream_t s)
{
  const Scenes sc = make_scenes(prim_off, nseg, n);
  const uint2* tl = static_cast<const uint2*>(treelets);
  // typical trees hold about 2 n / kTreelet treelets (plus one per small
  // scene); the loops are persistent, so any count is covered
  const uint32_t ga = grid_for(2 * n + uint64_t(sc.nseg) * kTreelet, kTreelet, 512);
  hipLaunchKernelGGL(k_fit_treelets, dim3(ga), dim3(kTreelet), 0, s, prims, kind, sc,
                     static_cast<Node*>(nodes), parent, leafprim, tl, treelet_count, queue,
                     static_cast<WideNode*>(wide));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  const uint32_t gb = grid_for(4 * n / kTreelet + 64, kBlock, 1024);
  hipLaunchKernelGGL(k_fit_top, dim3(gb), dim3(kBlock), 0, s, prims, kind, sc,
                     static_cast<Node*>(nodes), parent, leafprim, tl, treelet_count, flags,
                     static_cast<WideNode*>(wide));
  return hipGetLastError();
}

hipError_t
launch_enum_treelets(uint64_t n, const uint64_t* prim_off, uint32_t nseg, const uint32_t* parent,
                     const uint32_t* other_end, void* treelets, uint32_t* treelet_count,
                     hipStream_t s)
{
  const Scenes sc = make_scenes(prim_off, nseg, n);
  const uint64_t x = 2 * n - sc.nseg;
  const uint32_t g = uint32_t((x + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_enum_treelets, dim3(g), dim3(kBlock), 0, s, sc, parent, other_end,
                     static_cast<uint2*>(treelets), treelet_count);
  return hipGetLastError();
}

hipError_t
launch_fill_u32(uint32_t* p, uint64_t words, uint32_t value, hipStream_t s)
{
  if (words == 0)
    return hipSuccess;
  uint32_t g = grid_for(words, kBlock, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(kBlock), 0, s, p, words, value);
  return hipGetLastError();
}

} // namespace lbvh_eng
