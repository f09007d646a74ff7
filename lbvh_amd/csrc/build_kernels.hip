// gfx950 build kernels: centroid bounds, Morton encode (+ the radix sort's
// digit histograms), the radix tree with its boxes (one pass), and the
// bottom-up box fit of refits.  Each kernel cites the reference code whose
// results it reproduces bit for bit.
#include <cstdlib>

#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

constexpr int kBlock = 256;
constexpr int kCentroidMaxBlocks = 1024;
constexpr int kCentroidUnroll = 4;
constexpr int kEncodeMaxBlocks = 1024;
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
// reduction order cannot change the result for finite inputs (a NaN centre
// never replaces the running value: std_min(acc, NaN) keeps acc).  Centres
// are also written out (SoA) for the encode pass.  Each workgroup writes the
// union of its centres as one partial box; the encode kernel reduces the
// partials (no ticket, no drain: a workgroup exits as soon as its stores are
// issued).

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

// KIND: kPrimTriangle or kPrimAabb (a template argument, so that the loads
// of a thread's kCentroidUnroll primitives issue together).
template<int KIND>
__global__ __launch_bounds__(kBlock) void
k_centroid_bounds(const float* __restrict__ prims, uint64_t n,
                  float* __restrict__ cx, float* __restrict__ cy,
                  float* __restrict__ cz, float* __restrict__ partials,
                  uint32_t* __restrict__ zero_ptr, uint32_t zero_words)
{
  __shared__ float s_box[kBlock / kWave][6];
  const uint64_t tid = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;

  for (uint64_t j = tid; j < zero_words; j += stride)
    zero_ptr[j] = 0u;

  float lo[3] = { INFINITY, INFINITY, INFINITY };
  float hi[3] = { -INFINITY, -INFINITY, -INFINITY };
  // kCentroidUnroll primitives per thread per round, all loads issued first
  for (uint64_t i0 = tid; i0 < n; i0 += stride * kCentroidUnroll) {
    // unconditional loads (indexes past the end read the last primitive and
    // are not used), so that all of them are in flight together
    Box b[kCentroidUnroll];
#pragma unroll
    for (int u = 0; u < kCentroidUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      b[u] = load_prim_box(prims, KIND, i < n ? i : n - 1);
    }
#pragma unroll
    for (int u = 0; u < kCentroidUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      if (i >= n)
        continue;
      float c[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        float s = b[u].lo[k] + b[u].hi[k];
        c[k] = s * 0.5f;
        lo[k] = std_min(lo[k], c[k]);
        hi[k] = std_max(hi[k], c[k]);
      }
      cx[i] = c[0];
      cy[i] = c[1];
      cz[i] = c[2];
    }
  }
  block_reduce_box(lo, hi, s_box);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      partials[6 * blockIdx.x + k] = lo[k];
      partials[6 * blockIdx.x + 3 + k] = hi[k];
    }
  }
}

// ---- Morton encode --------------------------------------------------------------
// morton_curve_kernel (lbvh.h:1432-1473): scale = D / (extent + 1) with
// D = 1024 (morton_domain<4>) or 2^20 (morton_domain<8>); q = trunc((c - min)
// * scale); code = expand(qx) << 2 | expand(qy) << 1 | expand(qz) with the
// magic-mask expansions of morton_encoder<4>/<8> (lbvh.h:1352-1358,
// 1377-1385).  The 63-bit variant applies morton_encoder<8> to FP32 centres.
// Each workgroup also counts the digits of the radix passes first_pass ..
// passes - 1 (those whose global histogram the sort reads) in LDS and adds
// its nonzero counts to the global histogram with atomics.

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
// 1,024 threads: one encode workgroup per sort chunk leaves a single
// workgroup on most CUs at 1M keys, so it brings 16 waves to hide the load
// latency instead of 4.
constexpr int kEncodeBlock = 1024;

// counts != nullptr: workgroup g encodes the chunk [g * chunk, (g + 1) *
// chunk) and also writes its counts of pass first_pass to
// counts[d * gridDim.x + g] -- the chunk counts of the reduce-then-scan
// sort's first pass, which then needs no count pass of its own.
template<typename KeyT>
__global__ __launch_bounds__(kEncodeBlock) void
k_encode(const float* __restrict__ cx, const float* __restrict__ cy,
         const float* __restrict__ cz, uint64_t n,
         const float* __restrict__ partials, uint32_t nparts, float* __restrict__ bounds_out,
         KeyT* __restrict__ codes,
         uint32_t* __restrict__ hist, int digit_bits, int first_pass, int passes,
         uint32_t* __restrict__ counts, uint64_t chunk)
{
  extern __shared__ uint32_t s_hist[];
  __shared__ float s_pbox[kEncodeBlock / kWave][6];
  const int radix = 1 << digit_bits;
  const int words = radix * (passes - first_pass);
  for (int j = threadIdx.x; j < words; j += kEncodeBlock)
    s_hist[j] = 0u;
  // grid-stride over [0, n), or this workgroup's chunk
  uint64_t first = uint64_t(blockIdx.x) * kEncodeBlock + threadIdx.x;
  uint64_t limit = n;
  uint64_t stride = uint64_t(gridDim.x) * kEncodeBlock;
  if (counts) {
    const uint64_t b = uint64_t(blockIdx.x) * chunk;
    first = b + threadIdx.x;
    limit = (b + chunk < n) ? b + chunk : n;
    stride = kEncodeBlock;
  }
  // the centroid bounds: union of the bounds kernel's per-workgroup boxes
  // (finite or empty; every workgroup computes the same union)
  float blo[3] = { INFINITY, INFINITY, INFINITY };
  float bhi[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (uint32_t g = threadIdx.x; g < nparts; g += kEncodeBlock)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      blo[k] = std_min(blo[k], partials[6 * g + k]);
      bhi[k] = std_max(bhi[k], partials[6 * g + 3 + k]);
    }
  wave_reduce_box(blo, bhi);
  if (__lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_pbox[threadIdx.x / kWave][k] = blo[k];
      s_pbox[threadIdx.x / kWave][3 + k] = bhi[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; k++) {
    blo[k] = s_pbox[0][k];
    bhi[k] = s_pbox[0][3 + k];
  }
  for (int ww = 1; ww < kEncodeBlock / kWave; ww++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      blo[k] = std_min(blo[k], s_pbox[ww][k]);
      bhi[k] = std_max(bhi[k], s_pbox[ww][3 + k]);
    }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      bounds_out[k] = blo[k];
      bounds_out[3 + k] = bhi[k];
    }
  }
  const float dom = sizeof(KeyT) == 4 ? 1024.0f : 1048576.0f;
  float bmin[3], scale[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bmin[k] = blo[k];
    float ext = bhi[k] - bmin[k];
    float den = ext + 1.0f;
    scale[k] = dom / den;
  }
  const KeyT mask = KeyT(radix - 1);
  for (uint64_t i0 = first; i0 < limit; i0 += stride * kEncodeUnroll) {
    // unconditional loads (past the end: the last centre, unused), all in
    // flight together
    float x[kEncodeUnroll], y[kEncodeUnroll], z[kEncodeUnroll];
#pragma unroll
    for (int u = 0; u < kEncodeUnroll; u++) {
      const uint64_t i = i0 + u * stride;
      const uint64_t j = i < limit ? i : limit - 1;
      x[u] = cx[j];
      y[u] = cy[j];
      z[u] = cz[j];
    }
#pragma unroll
    for (int u = 0; u < kEncodeUnroll; u++) {
      uint64_t i = i0 + u * stride;
      if (i >= limit)
        continue;
      float dx = x[u] - bmin[0];
      float dy = y[u] - bmin[1];
      float dz = z[u] - bmin[2];
      float qx = dx * scale[0];
      float qy = dy * scale[1];
      float qz = dz * scale[2];
      KeyT code = encode<KeyT>(qx, qy, qz);
      codes[i] = code;
      for (int p = first_pass; p < passes; p++) {
        uint32_t d = uint32_t((code >> (p * digit_bits)) & mask);
        atomicAdd(&s_hist[(p - first_pass) * radix + d], 1u);
      }
    }
  }
  __syncthreads();
  // the workgroup's counts go straight into the global histogram (zeroed
  // by k_centroid_bounds): no per-workgroup rows and no reduction kernel
  for (int j = threadIdx.x; j < words; j += kEncodeBlock) {
    const uint32_t v = s_hist[j];
    if (v)
      atomicAdd(&hist[j], v);
  }
  if (counts)
    for (int d = threadIdx.x; d < radix; d += kEncodeBlock)
      counts[size_t(d) * gridDim.x + blockIdx.x] = s_hist[d];
}

// ---- prefix lengths of sorted codes ----------------------------------------------
// calc_delta (lbvh.h:1602-1621): delta(j, k) = clz(code_j ^ code_k) when the
// codes differ, else clz(j ^ k) + bits(code), with scene-local positions.

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

// Union of element boxes [a, b] left to right, from a bottom-up segment
// tree over m elements (leaves at bt[m + e]; bt[p] = union(bt[2p],
// bt[2p + 1])).  std::min / std::max with the left operand kept on ties
// are associative on non-NaN values, so any order-preserving grouping gives
// the bits of the reference's union of its two children, recursively.
__device__ __forceinline__ Box
bt_query(const Box* __restrict__ bt, int m, int a, int b)
{
  int l = a + m, r = b + m + 1;
  Box rl, rr;
  bool hl = false, hr = false;
  while (l < r) {
    if (l & 1) {
      rl = hl ? box_union(rl, bt[l]) : bt[l];
      hl = true;
      l++;
    }
    if (r & 1) {
      r--;
      rr = hr ? box_union(bt[r], rr) : bt[r];
      hr = true;
    }
    l >>= 1;
    r >>= 1;
  }
  if (hl && hr)
    return box_union(rl, rr);
  return hl ? rl : rr;
}

__device__ __forceinline__ bool
box_has_nan(const Box& b)
{
  bool x = false;
#pragma unroll
  for (int q = 0; q < 3; q++)
    x = x || (b.lo[q] != b.lo[q]) || (b.hi[q] != b.hi[q]);
  return x;
}

// ---- bottom-up fit -------------------------------------------------------------
// builder::fit_boxes (lbvh.h:1827-1866) computes every internal box as
// union(left box, right box), left operand first, where a leaf's box is the
// converter applied to the primitive its index names.  In the Karras
// layout the children of a node with split g are g and g + 1 (leaf
// positions or node indices), so in the top tree the parent link's side and
// sibling-is-leaf bits locate the sibling without reading the node, and the
// second of the two child computations to finish computes the parent's box.
//
// Pass 1 (k_fit_treelets): each treelet -- a subtree of at most kTreelet
// leaves under a top node -- is fitted by one workgroup in LDS from a
// segment tree over its leaf boxes.  Workgroups pull treelets from the list
// k_enum_treelets wrote.
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

// One workgroup per treelet (pulled from the list): the leaf boxes go to LDS
// as the leaves of a segment tree, and every internal node of the treelet --
// node indices [lo, hi - 1] when its root is a right child (numbered lo),
// [lo + 1, hi] when a left child (numbered hi) -- takes its children's boxes
// as the unions over their leaf ranges [L, g] and [g + 1, R] (std::min /
// std::max keeping the left operand are associative on non-NaN values, so
// the bits equal the reference's recursion).  A treelet whose leaf boxes hold
// a NaN is fitted in tree order instead, in passes, each node once both
// children are done.  The node's range comes from other_end; its split g
// from its children (node g or leaf g on the left, node g + 1 or leaf g + 1
// on the right).
__global__ __launch_bounds__(kTreelet) void
k_fit_treelets(const float* __restrict__ prims_all, int kind, Scenes sc, Node* __restrict__ nodes_all,
               const uint32_t* __restrict__ other_end_all,
               const uint32_t* __restrict__ leafprim_all, const uint2* __restrict__ treelets,
               const uint32_t* __restrict__ count_ptr, uint32_t* __restrict__ queue)
{
  // leaf boxes at s_bt[m + e]; below m the segment tree, or (NaN) node
  // base + j's box at s_bt[1 + j]
  __shared__ Box s_bt[2 * kTreelet];
  __shared__ uint8_t s_done[kTreelet];
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
    const uint2 en = treelets[item];
    const TreeletScene ts = treelet_scene(sc, en);
    const int64_t lo = ts.lo, hi = ts.hi, n = ts.n;
    const int m = int(hi - lo + 1);
    if (lo >= 0 && m >= 2 && m <= kTreelet && hi < n) {
      const float* prims = prims_all + prim_stride(kind) * ts.P;
      Node* nodes = nodes_all + ts.nb;
      const uint32_t* other_end = other_end_all + ts.nb;
      const uint32_t* leafprim = leafprim_all ? leafprim_all + ts.P : nullptr;
      // the root is numbered lo when it is a right child (and for a
      // whole-tree treelet, whose root is node 0)
      const bool root_lo = (en.y & 0x80000000u) || (lo == 0 && hi == n - 1);
      const int64_t base = root_lo ? lo : lo + 1;
      bool bad = false;
      if (t < m)
        bad = box_has_nan(s_bt[m + t] = leaf_box(prims, kind, leafprim, lo + t));
      const bool nan = __syncthreads_or(bad ? 1 : 0) != 0;
      // this thread's internal node k = base + t (t < m - 1): range, children, split
      const int64_t k = base + t;
      bool mine = false;
      int64_t L = 0, R = 0, g = 0;
      uint32_t lv = 0, rv = 0;
      if (t < m - 1 && k < n - 1) {
        const int64_t oe = other_end[k];
        L = k < oe ? k : oe;
        R = k < oe ? oe : k;
        const uint2 ch = *reinterpret_cast<const uint2*>(&nodes[k].left);
        lv = ch.x;
        rv = ch.y;
        g = !(lv & kLeafBit) ? int64_t(lv) : (!(rv & kLeafBit) ? int64_t(rv) - 1 : L);
        // ranges and children come from the tree kernel; never false for a valid tree
        mine = L >= lo && R <= hi && L <= g && g < R;
        s_done[t] = 0;
      }
      __syncthreads();
      if (!nan) {
        // the segment tree over the m leaf boxes
        int leaf0 = 1;
        while (leaf0 < m)
          leaf0 <<= 1;
        for (int half = leaf0 / 2; half >= 1; half >>= 1) {
          for (int q = t; q < half; q += kTreelet) {
            const int p = half + q;
            if (p < m)
              s_bt[p] = box_union(s_bt[2 * p], s_bt[2 * p + 1]);
          }
          __syncthreads();
        }
        if (mine) {
          const Box lb = bt_query(s_bt, m, int(L - lo), int(g - lo));
          const Box rb = bt_query(s_bt, m, int(g + 1 - lo), int(R - lo));
          store_node(&nodes[k], box_union(lb, rb), lv, rv);
        }
      } else {
        // children before parents: node box of base + j at s_bt[1 + j]
        for (;;) {
          int changed = 0;
          if (mine && !s_done[t]) {
            const bool l_leaf = (lv & kLeafBit) != 0, r_leaf = (rv & kLeafBit) != 0;
            const int lj = int(g - base), rj = int(g + 1 - base);
            const bool l_ok = l_leaf || (lj >= 0 && lj < m - 1 && s_done[lj]);
            const bool r_ok = r_leaf || (rj >= 0 && rj < m - 1 && s_done[rj]);
            if (l_ok && r_ok) {
              const Box lb = l_leaf ? s_bt[m + int(g - lo)] : s_bt[1 + lj];
              const Box rb = r_leaf ? s_bt[m + int(g + 1 - lo)] : s_bt[1 + rj];
              const Box b = box_union(lb, rb);
              s_bt[1 + t] = b;
              store_node(&nodes[k], b, lv, rv);
              changed = 1;
            }
          }
          __syncthreads();
          if (changed)
            s_done[t] = 1;
          if (!__syncthreads_or(changed))
            break;
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
          uint32_t* __restrict__ flags_all)
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

// ---- bottom-up construction and fit in one pass ---------------------------------
// The Karras tree of sorted codes is the binary radix tree of the keys
// (code, position): node ranges and splits depend only on the adjacent
// prefix lengths delta(x) = delta(x, x + 1) of calc_delta (lbvh.h:1602-1621),
// with delta = -1 past either end of a scene.  A node [l, r] is the left
// child of the node split at r when delta(r) > delta(l - 1), else the right
// child of the node split at l - 1 (the two are never equal); a left child
// is numbered r and a right child l, the root 0 -- the numbering of
// builder_kernel (lbvh.h:1713-1731).  So the tree can be grown from the
// leaves (Apetrei, "Fast and Simple Agglomerative LBVH Construction", 2014):
// the second child to reach a split completes that node -- its range, its
// box union(left, right) as fit_boxes computes it (lbvh.h:1827-1866), its
// number, its node record and its children's parent links.
//
// k_tree runs this one level of workgroups after another, in LDS:
//   level 0: a workgroup per kTreeLeaves sorted positions; its elements are
//     the leaves;
//   level k + 1: the last of each kTreeFan level-k workgroups to finish
//     (an agent-scope ticket; nobody waits) takes their listed subtrees, in
//     position order, as its elements.  When they hold more than the
//     workgroup's slots, it merges them kTreeSubFan lists at a time into
//     lists of its own (each such merge is an ordinary step of the same
//     algorithm over a shorter position range) and then merges those.
// Within a workgroup, with d[e] the delta left of element e, the node split
// between elements j and j + 1 spans elements [a, b] where d[a] and d[b + 1]
// are the nearest deltas below d[j + 1] on either side (searched in a min
// tree over d); it lies in the workgroup when both exist.  Its children are
// the elements j or j + 1 or the nodes numbered j and j + 1 in element
// space, so boxes follow in passes, each node once both
// children are done (the reference's union order).  Nodes and elements
// whose parent does not lie in the workgroup -- at most two per tree level,
// so at most kTreeSegCap -- are listed in position order
// ({l, r, delta(l - 1), delta(r), box}).  The single workgroup of the last
// level completes every remaining node.  Segments and tickets cross
// workgroups (and XCDs) through coherent stores, a vector-memory drain and
// the ticket atomic; tickets reset themselves.
//
// Outputs: node records, parent links (kLink*), other_end,
// leafprim, and the fit flags left at 0 (what a refit starts from).

constexpr int kTreeLeaves = 512;                        // level-0 positions per workgroup
constexpr int kTreeFan = 64;                            // workgroups merged per level
constexpr int kTreeSubFan = 8;                          // lists per merge that always fits
constexpr int kTreeSubLists = kTreeFan / kTreeSubFan;   // lists of a split merge
constexpr int kTreeSegCap = 192;                        // 2 x (tree height bound + 1)
constexpr int kTreeSlots = kTreeSubFan * kTreeSegCap;   // elements of one workgroup
static_assert(kTreeSlots >= kTreeLeaves, "level-0 slots");
static_assert(kTreeFan % kTreeSubFan == 0 && kTreeSubLists * kTreeSegCap <= kTreeSlots,
              "split merges fit");
static_assert(kTreeFan <= kWave, "one wave scans a group's list counts");

// A listed subtree: leaf range [l, r], delta(l - 1) and delta(r) (packed as
// int16 pairs), box.
struct TreeEntry
{
  uint32_t l, r;
  uint32_t d; // (delta(l - 1) & 0xffff) | delta(r) << 16
  float box[6];
};
static_assert(sizeof(TreeEntry) == 36, "TreeEntry layout");

// delta(x, x + 1) over sorted batch positions; -1 at scene edges.
template<typename KeyT, typename CodeT, bool SEG>
__device__ __forceinline__ int
adj_delta(const KeyT* __restrict__ keys, const uint64_t* __restrict__ prim_off, int64_t N,
          int64_t x)
{
  if (x < 0 || x >= N - 1)
    return -1;
  const KeyT a = keys[x];
  const KeyT b = keys[x + 1];
  if constexpr (SEG) {
    if ((a >> 30) != (b >> 30))
      return -1;
  }
  const CodeT ca = key_code<KeyT, CodeT, SEG>(a);
  const CodeT cb = key_code<KeyT, CodeT, SEG>(b);
  if (ca != cb)
    return CodeTraits<CodeT>::clz(ca ^ cb);
  int64_t i = x;
  if constexpr (SEG)
    i -= int64_t(prim_off[uint32_t(a >> 30)]);
  return CodeTraits<CodeT>::clz(CodeT(uint64_t(i ^ (i + 1)))) + CodeTraits<CodeT>::bits;
}

// One scene's arrays, located from any of its sorted positions.
struct TreeScene
{
  int64_t P, n;
  const float* prims;
  Node* nodes;
  uint32_t* parent;
  uint32_t* leafprim;
  uint32_t* flags;
  uint32_t* other_end;
};

template<typename KeyT, bool SEG>
__device__ __forceinline__ TreeScene
tree_scene(const KeyT* __restrict__ keys, const Scenes& sc, int64_t x, const float* prims_all,
           int kind, Node* nodes_all, uint32_t* parent_all,
           uint32_t* leafprim_all, uint32_t* flags_all, uint32_t* other_end_all)
{
  uint32_t seg = 0;
  int64_t P = 0, n = sc.n;
  if constexpr (SEG) {
    seg = uint32_t(keys[x] >> 30);
    P = int64_t(sc.prim_off[seg]);
    n = int64_t(sc.prim_off[seg + 1]) - P;
  }
  const int64_t nb = P - int64_t(seg);
  const int64_t lb = 2 * P - int64_t(seg);
  TreeScene t;
  t.P = P;
  t.n = n;
  t.prims = prims_all + int64_t(prim_stride(kind)) * P;
  t.nodes = nodes_all + nb;
  t.parent = parent_all + lb;
  t.leafprim = leafprim_all ? leafprim_all + P : nullptr;
  t.flags = flags_all + nb;
  t.other_end = other_end_all + nb;
  return t;
}

// Child value of the leaf at sorted position x.
__device__ __forceinline__ uint32_t
tree_leaf_value(const TreeScene& ts, const uint32_t* __restrict__ ids, int64_t x)
{
  return (ts.leafprim ? uint32_t(int64_t(ids[x]) - ts.P) : uint32_t(x - ts.P)) | kLeafBit;
}

struct TreeLevels
{
  int levels;
  uint32_t slot_cap;        // elements merged at once (<= kTreeSlots), else split
  uint32_t nblocks[16];     // workgroups of each level (the last has 1)
  uint32_t seg_base[16];    // first segment of each level
  uint32_t tmp_base[16];    // first of each level's split-merge segments (kTreeSubLists each)
  uint32_t ticket_base[16]; // first ticket of each level's groups
};

constexpr int kTreeLeaf0 = 2048; // leaves of the delta min-tree (>= kTreeSlots + 1)
static_assert(kTreeLeaf0 >= kTreeSlots + 1, "delta tree size");
constexpr int kTreeLeaf0First = 1024; // the same at level 0 (>= kTreeLeaves + 1)
static_assert(kTreeLeaf0First >= kTreeLeaves + 1, "level-0 delta tree size");
constexpr int kTreeUpThreads = 1024; // workgroup of the upper levels (one CU each: LDS-bound)
constexpr int16_t kTreeInf = 0x7fff;

// Largest i <= j with d[i] < x, or -1 (tree: heap-ordered minima of d,
// leaves at leaf0 + i).
__device__ __forceinline__ int
tree_left_less(const int16_t* __restrict__ tree, int leaf0, int j, int x)
{
  int p = j + leaf0;
  if (tree[p] < x)
    return j;
  while (p > 1) {
    if ((p & 1) && tree[p - 1] < x) {
      p = p - 1;
      while (p < leaf0)
        p = (tree[2 * p + 1] < x) ? 2 * p + 1 : 2 * p;
      return p - leaf0;
    }
    p >>= 1;
  }
  return -1;
}

// Smallest i >= j with d[i] < x, or -1.
__device__ __forceinline__ int
tree_right_less(const int16_t* __restrict__ tree, int leaf0, int j, int x)
{
  int p = j + leaf0;
  if (tree[p] < x)
    return j;
  while (p > 1) {
    if (!(p & 1) && tree[p + 1] < x) {
      p = p + 1;
      while (p < leaf0)
        p = (tree[2 * p] < x) ? 2 * p : 2 * p + 1;
      return p - leaf0;
    }
    p >>= 1;
  }
  return -1;
}

// FIRST: the level-0 kernel (leaves, 512 elements, ~40 KB of LDS: four workgroups fit a CU);
// else the kernel of the levels above (its workgroup b starts as level-1
// workgroup b).
template<typename KeyT, typename CodeT, bool SEG, bool FIRST>
__global__ __launch_bounds__(FIRST ? kTreeLeaves : kTreeUpThreads) void
k_tree(const KeyT* __restrict__ keys, const uint32_t* __restrict__ ids, Scenes sc,
       const float* __restrict__ prims_all, int kind, Node* __restrict__ nodes_all,
       uint32_t* __restrict__ parent_all,
       uint32_t* __restrict__ leafprim_all, uint32_t* __restrict__ flags_all,
       uint32_t* __restrict__ other_end_all, TreeLevels lv, uint32_t* __restrict__ seg_count,
       TreeEntry* __restrict__ seg_entries, uint32_t* __restrict__ tickets)
{
  constexpr int SL = FIRST ? kTreeLeaves : kTreeSlots;      // elements per workgroup
  constexpr int NT = FIRST ? kTreeLeaves : kTreeUpThreads;  // threads
  constexpr int EPT = (SL + NT - 1) / NT;                   // elements per thread (lists)
  // elements (level 0: leaves; then listed subtrees), in position order
  __shared__ uint32_t s_el_l[SL];
  __shared__ uint32_t s_el_r[SL];
  // element boxes at s_bt[m + e]; below m: the segment tree over them, or
  // (NaN boxes present) node j's box at s_bt[1 + j]
  __shared__ Box s_bt[2 * SL];
  // scene of each element (batches above level 0; level 0 reads it from the key)
  __shared__ uint32_t s_el_scene[(SEG && !FIRST) ? SL : 1];
  __shared__ uint32_t s_el_val[SL]; // child value when the element is one leaf
  // min tree over d (leaves at kTreeLeaf0): d[e] is the delta left of
  // element e, d[m] the delta right of the last
  __shared__ int16_t s_tree[2 * (FIRST ? kTreeLeaf0First : kTreeLeaf0)];
  __shared__ int16_t s_dsep[FIRST ? 1 : SL + 1];
  int16_t* const s_d = FIRST ? s_tree + kTreeLeaf0First : s_dsep; // level 0: the tree's own leaves
  // split j (between elements j and j + 1): its node's element range [a, b]
  __shared__ uint16_t s_na[SL];
  __shared__ uint16_t s_nb[SL];
  __shared__ uint8_t s_inb[SL];     // 1: the node's range lies in the workgroup's
  __shared__ uint8_t s_round[SL];   // pass in which its box was completed (0: not yet)
  __shared__ uint16_t s_kid[SL];    // element-space node number -> split
  __shared__ uint16_t s_esc[SL];    // listed subtree starting at element e (+1; 0: none)
  __shared__ uint32_t s_wsum[NT / kWave];
  __shared__ uint32_t s_pre[kWave + 1];
  __shared__ uint32_t s_last;

  const int t = threadIdx.x;
  const int lane = __lane_id();
  const int w = t / kWave;
  const int64_t N = sc.n;
  int level = FIRST ? 0 : 1;
  uint32_t blk = blockIdx.x;
  int m = 0;

  // Elements from the lists of segments [first, first + nl), in position
  // order; false (nothing loaded) when they number more than cap.
  auto load_lists = [&](uint32_t first, uint32_t nl, uint32_t cap) -> bool {
    if (t < kWave) {
      const uint32_t cnt = uint32_t(t) < nl ? coherent_load_u32(&seg_count[first + t]) : 0u;
      uint32_t incl = cnt;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, kWave);
        if (lane >= off)
          incl += y;
      }
      if (uint32_t(t) < nl)
        s_pre[t + 1] = incl;
      if (t == 0)
        s_pre[0] = 0;
    }
    __syncthreads();
    const uint32_t total = s_pre[nl];
    if (total > cap) {
      __syncthreads(); // every thread has read s_pre before it is rewritten
      return false;
    }
    m = int(total);
    for (int e = t; e < m; e += NT) {
      uint32_t c = 0, chi = nl; // the last list starting at or before e
      while (chi - c > 1) {
        const uint32_t mid = (c + chi) >> 1;
        if (s_pre[mid] <= uint32_t(e))
          c = mid;
        else
          chi = mid;
      }
      const TreeEntry* en = &seg_entries[size_t(first + c) * kTreeSegCap + (uint32_t(e) - s_pre[c])];
      const uint32_t el = coherent_load_u32(&en->l);
      const uint32_t er = coherent_load_u32(&en->r);
      const uint32_t d = coherent_load_u32(&en->d);
      Box bx;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        bx.lo[q] = coherent_load_f32(&en->box[q]);
        bx.hi[q] = coherent_load_f32(&en->box[3 + q]);
      }
      s_el_l[e] = el;
      s_el_r[e] = er;
      s_bt[m + e] = bx;
      s_d[e + 1] = int16_t(d >> 16);
      if (e == 0)
        s_d[0] = int16_t(d & 0xffffu);
      int64_t P = 0;
      if constexpr (SEG) {
        const uint32_t sg = uint32_t(keys[el] >> 30);
        P = int64_t(sc.prim_off[sg]);
        if constexpr (!FIRST)
          s_el_scene[e] = sg;
      }
      if (el == er)
        s_el_val[e] = (leafprim_all ? uint32_t(int64_t(ids[el]) - P) : uint32_t(int64_t(el) - P)) |
                      kLeafBit;
    }
    __syncthreads();
    return true;
  };

  // Group blk of level `level`: the lists of its level-(level - 1) members,
  // all at once when they fit, else the first kTreeSubFan of them (a split
  // merge: sub counts the merges done so far).
  uint32_t out_seg = 0, sub = 0, nsub = 0, grp_first = 0, grp_members = 0;
  auto begin_group = [&]() {
    grp_first = lv.seg_base[level - 1] + blk * kTreeFan;
    grp_members = (lv.nblocks[level - 1] - blk * kTreeFan) < uint32_t(kTreeFan)
                    ? lv.nblocks[level - 1] - blk * kTreeFan
                    : uint32_t(kTreeFan);
    if (load_lists(grp_first, grp_members, lv.slot_cap)) {
      nsub = 0;
      out_seg = lv.seg_base[level] + blk;
      return;
    }
    nsub = (grp_members + kTreeSubFan - 1) / kTreeSubFan;
    sub = 0;
    load_lists(grp_first, grp_members < uint32_t(kTreeSubFan) ? grp_members : uint32_t(kTreeSubFan),
               uint32_t(kTreeSlots));
    out_seg = lv.tmp_base[level] + blk * kTreeSubLists;
  };

  if constexpr (FIRST) {
    // ---- level 0 elements: the leaves -------------------------------------------
    const int64_t lo = int64_t(blk) * kTreeLeaves;
    const int64_t hi = (N - lo < kTreeLeaves) ? N : lo + kTreeLeaves;
    m = int(hi - lo);
    s_d[t] = int16_t(adj_delta<KeyT, CodeT, SEG>(keys, sc.prim_off, N, lo - 1 + t));
    if (t == 0)
      s_d[kTreeLeaves] =
        int16_t(adj_delta<KeyT, CodeT, SEG>(keys, sc.prim_off, N, lo - 1 + kTreeLeaves));
    if (t < m) {
      const int64_t x = lo + t;
      const TreeScene ts = tree_scene<KeyT, SEG>(keys, sc, x, prims_all, kind, nodes_all,
                                                 parent_all, leafprim_all, flags_all,
                                                 other_end_all);
      const int64_t lx = x - ts.P;
      if (lx < ts.n - 1)
        ts.flags[lx] = 0u; // fit flags (refits) start at 0
      uint32_t prim = uint32_t(lx);
      if (ts.leafprim) {
        prim = uint32_t(int64_t(ids[x]) - ts.P);
        ts.leafprim[lx] = prim;
      }
      s_el_l[t] = uint32_t(x);
      s_el_r[t] = uint32_t(x);
      s_el_val[t] = prim | kLeafBit;
      s_bt[m + t] = load_prim_box(ts.prims, kind, uint64_t(prim));
    }
    __syncthreads();
  } else {
    begin_group();
  }

  for (;;) {
    // ---- nodes of this workgroup: ranges from the nearest smaller deltas -------
    // level 0: kTreeLeaf0First leaves, d already in place; above: the next power of two
    int leaf0 = kTreeLeaf0First;
    if constexpr (!FIRST) {
      leaf0 = 2;
      while (leaf0 < m + 1)
        leaf0 <<= 1;
      for (int q = t; q <= m; q += NT)
        s_tree[leaf0 + q] = s_d[q];
    }
    for (int q = m + 1 + t; q < leaf0; q += NT)
      s_tree[leaf0 + q] = kTreeInf;
    __syncthreads();
    // exact NaN semantics need the unions in tree order (passes below)
    int any_nan = 0;
    for (int e = t; e < m; e += NT)
      any_nan |= box_has_nan(s_bt[m + e]) ? 1 : 0;
    const bool nan = __syncthreads_or(any_nan) != 0;
    for (int half = leaf0 / 2; half >= 1; half >>= 1) {
      for (int q = t; q < half; q += NT) {
        const int p = half + q;
        const int16_t a0 = s_tree[2 * p], a1 = s_tree[2 * p + 1];
        s_tree[p] = a0 < a1 ? a0 : a1;
        if (!nan && p < m)
          s_bt[p] = box_union(s_bt[2 * p], s_bt[2 * p + 1]);
      }
      __syncthreads();
    }
    for (int j = t; j < m - 1; j += NT) {
      const int x = s_d[j + 1];
      uint8_t inb = 0;
      if (x >= 0) {
        const int i = tree_left_less(s_tree, leaf0, j, x);
        const int k = tree_right_less(s_tree, leaf0, j + 2, x);
        if (i >= 0 && k >= 0) {
          inb = 1;
          s_na[j] = uint16_t(i);
          s_nb[j] = uint16_t(k - 1);
        }
      }
      s_inb[j] = inb;
      s_round[j] = 0;
    }
    for (int e = t; e < m; e += NT) {
      s_esc[e] = 0;
      s_kid[e] = 0xffff;
    }
    __syncthreads();
    for (int j = t; j < m - 1; j += NT) {
      if (s_inb[j]) {
        const int a = s_na[j], b = s_nb[j];
        s_kid[(s_d[b + 1] > s_d[a]) ? b : a] = uint16_t(j);
      }
    }
    __syncthreads();
    // ---- boxes, children before parents (LDS only: a barrier waits for every
    // outstanding global store, so the records are written afterwards) ------
    for (int pass = 1; nan; pass++) {
      int changed = 0;
      for (int j = t; j < m - 1; j += NT) {
        if (!s_inb[j] || s_round[j] != 0)
          continue;
        const int a = s_na[j], b = s_nb[j];
        const bool l_el = (a == j);
        const bool r_el = (j + 1 == b);
        int lc = -1, rc = -1;
        if (!l_el) {
          lc = s_kid[j];
          const int rd = s_round[lc];
          if (rd == 0 || rd >= pass)
            continue;
        }
        if (!r_el) {
          rc = s_kid[j + 1];
          const int rd = s_round[rc];
          if (rd == 0 || rd >= pass)
            continue;
        }
        const Box lb = l_el ? s_bt[m + j] : s_bt[1 + lc];
        const Box rb = r_el ? s_bt[m + j + 1] : s_bt[1 + rc];
        s_bt[1 + j] = box_union(lb, rb);
        s_round[j] = uint8_t(pass);
        changed = 1;
      }
      if (!__syncthreads_or(changed))
        break;
    }
    // ---- records of the completed nodes ------------------------------------------
    // children's parent links by split (consecutive positions), node records
    // by node number (consecutive records)
    auto node_scene = [&](int j, int64_t& P, int64_t& n, uint32_t& seg) {
      P = 0;
      n = N;
      seg = 0;
      if constexpr (SEG) {
        if constexpr (FIRST)
          seg = uint32_t(keys[s_el_l[j]] >> 30);
        else
          seg = s_el_scene[j];
        P = int64_t(sc.prim_off[seg]);
        n = int64_t(sc.prim_off[seg + 1]) - P;
      }
    };
    for (int j = t; j < m - 1; j += NT) {
      if (!s_inb[j])
        continue;
      const int a = s_na[j], b = s_nb[j];
      const int64_t L = s_el_l[a];
      const int64_t R = s_el_r[b];
      const int64_t g = s_el_r[j];
      int64_t P, n;
      uint32_t seg;
      node_scene(j, P, n, seg);
      const int64_t k = (int(s_d[b + 1]) > int(s_d[a])) ? R - P : L - P;
      const bool lleaf = (a == j) && s_el_l[j] == s_el_r[j];
      const bool rleaf = (j + 1 == b) && s_el_l[j + 1] == s_el_r[j + 1];
      const int64_t lg = g - P;
      const uint32_t up = uint32_t(k) | ((R - L + 1) > kTreelet ? kLinkGlobal : 0u);
      uint32_t* parent = parent_all + (2 * P - int64_t(seg));
      parent[lleaf ? (n - 1) + lg : lg] = up | (rleaf ? kLinkSibLeaf : 0u);
      parent[rleaf ? (n - 1) + lg + 1 : lg + 1] = up | kLinkRight | (lleaf ? kLinkSibLeaf : 0u);
    }
    for (int e = t; e < m; e += NT) {
      const uint32_t jj = s_kid[e];
      if (jj == 0xffffu)
        continue; // no node is numbered e
      const int j = int(jj);
      const int a = s_na[j], b = s_nb[j];
      const bool l_el = (a == j);
      const bool r_el = (j + 1 == b);
      Box lb, rb, bx;
      if (nan) {
        lb = l_el ? s_bt[m + j] : s_bt[1 + s_kid[j]];
        rb = r_el ? s_bt[m + j + 1] : s_bt[1 + s_kid[j + 1]];
        bx = s_bt[1 + j];
      } else {
        lb = bt_query(s_bt, m, a, j);
        rb = bt_query(s_bt, m, j + 1, b);
        bx = box_union(lb, rb);
      }
      const int64_t L = s_el_l[a];
      const int64_t R = s_el_r[b];
      const int64_t g = s_el_r[j];
      int64_t P, n;
      uint32_t seg;
      node_scene(j, P, n, seg);
      const int64_t nb = P - int64_t(seg);
      const int64_t k = (int(s_d[b + 1]) > int(s_d[a])) ? R - P : L - P;
      const bool lleaf = l_el && s_el_l[j] == s_el_r[j];
      const bool rleaf = r_el && s_el_l[j + 1] == s_el_r[j + 1];
      const int64_t lg = g - P;
      const uint32_t lvv = lleaf ? s_el_val[j] : uint32_t(lg);
      const uint32_t rvv = rleaf ? s_el_val[j + 1] : uint32_t(lg + 1);
      store_node(&nodes_all[nb + k], bx, lvv, rvv);
      other_end_all[nb + k] = uint32_t((k == L - P) ? R - P : L - P);
    }
    // ---- listed subtrees: elements and nodes whose parent is not in here ------
    auto parent_out = [&](int a, int b) {
      const int da = s_d[a], db = s_d[b + 1];
      if (da < 0 && db < 0)
        return false; // the scene's root
      const int ps = (db > da) ? b : a - 1;
      return ps < 0 || ps > m - 2 || !s_inb[ps];
    };
    for (int e = t; e < m; e += NT)
      if (parent_out(e, e))
        s_esc[e] = uint16_t(e + 1); // an element stands alone
    __syncthreads();
    for (int j = t; j < m - 1; j += NT)
      if (s_inb[j] && parent_out(s_na[j], s_nb[j]))
        s_esc[s_na[j]] = uint16_t(kTreeSlots + 1 + j); // a completed node
    __syncthreads();
    // position order of the listed subtrees: exclusive count of list starts
    // (thread t holds elements [t * EPT, t * EPT + EPT))
    uint32_t mine = 0;
#pragma unroll
    for (int h = 0; h < EPT; h++) {
      const int e = t * EPT + h;
      mine += (e < m && s_esc[e]) ? 1u : 0u;
    }
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, kWave);
      if (lane >= off)
        incl += y;
    }
    if (lane == kWave - 1)
      s_wsum[w] = incl;
    __syncthreads();
    uint32_t base = incl - mine;
    uint32_t total = 0;
#pragma unroll
    for (int q = 0; q < NT / kWave; q++) {
      const uint32_t v = s_wsum[q];
      base += (q < w) ? v : 0u;
      total += v;
    }
    TreeEntry* out = &seg_entries[size_t(FIRST ? blk : out_seg) * kTreeSegCap];
    uint32_t o = base;
#pragma unroll
    for (int h = 0; h < EPT; h++) {
      const int e = t * EPT + h;
      if (e >= m || !s_esc[e])
        continue;
      if (o >= uint32_t(kTreeSegCap))
        break; // cannot happen: two per tree level at most
      const uint32_t tag = s_esc[e];
      int a, b;
      Box bx;
      if (tag > uint32_t(kTreeSlots)) {
        const int j = int(tag) - kTreeSlots - 1;
        a = s_na[j];
        b = s_nb[j];
        bx = nan ? s_bt[1 + j] : bt_query(s_bt, m, a, b);
      } else {
        a = b = e;
        bx = s_bt[m + e];
      }
      TreeEntry* en = &out[o++];
      coherent_store_u32(&en->l, s_el_l[a]);
      coherent_store_u32(&en->r, s_el_r[b]);
      coherent_store_u32(&en->d, (uint32_t(int(s_d[a])) & 0xffffu) | (uint32_t(int(s_d[b + 1])) << 16));
#pragma unroll
      for (int q = 0; q < 3; q++) {
        coherent_store_f32(&en->box[q], bx.lo[q]);
        coherent_store_f32(&en->box[3 + q], bx.hi[q]);
      }
    }
    if (t == 0)
      coherent_store_u32(&seg_count[FIRST ? blk : out_seg],
                         total < uint32_t(kTreeSegCap) ? total : uint32_t(kTreeSegCap));
    if constexpr (FIRST)
      return; // level 0 hands over to the next kernel
    if (nsub) {
      // ---- split merge: the next kTreeSubFan lists, then this workgroup's own ----
      vmem_drain();
      __syncthreads();
      const uint32_t tmp0 = lv.tmp_base[level] + blk * kTreeSubLists;
      if (++sub < nsub) {
        const uint32_t rest = grp_members - sub * kTreeSubFan;
        load_lists(grp_first + sub * kTreeSubFan,
                   rest < uint32_t(kTreeSubFan) ? rest : uint32_t(kTreeSubFan),
                   uint32_t(kTreeSlots));
        out_seg = tmp0 + sub;
      } else {
        load_lists(tmp0, nsub, uint32_t(kTreeSlots));
        out_seg = lv.seg_base[level] + blk;
        nsub = 0;
      }
      continue;
    }
    if (level + 1 >= lv.levels)
      return; // the last level completes every node
    // ---- hand over: the last of this group's workgroups goes on ---------------
    vmem_drain();
    __syncthreads();
    const uint32_t grp = blk / kTreeFan;
    const uint32_t members = (lv.nblocks[level] - grp * kTreeFan) < uint32_t(kTreeFan)
                               ? lv.nblocks[level] - grp * kTreeFan
                               : uint32_t(kTreeFan);
    if (t == 0) {
      uint32_t* tk = &tickets[lv.ticket_base[level] + grp];
      const uint32_t old = atomic_add_agent(tk, 1u);
      const bool last = old + 1u == members;
      if (last)
        coherent_store_u32(tk, 0u); // reset for the next build
      s_last = last ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last)
      return;
    level++;
    blk = grp;
    begin_group();
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
  return grid_for(n, uint64_t(kEncodeBlock) * 4, kEncodeMaxBlocks);
}

hipError_t
launch_centroid_bounds(const float* prims, int kind, uint64_t n, float* cx, float* cy,
                       float* cz, float* partials, uint32_t* zero_ptr, uint32_t zero_words,
                       hipStream_t s)
{
  uint32_t g = centroid_partials_needed(n);
  if (kind == kPrimTriangle)
    hipLaunchKernelGGL(k_centroid_bounds<kPrimTriangle>, dim3(g), dim3(kBlock), 0, s, prims, n,
                       cx, cy, cz, partials, zero_ptr, zero_words);
  else
    hipLaunchKernelGGL(k_centroid_bounds<kPrimAabb>, dim3(g), dim3(kBlock), 0, s, prims, n, cx,
                       cy, cz, partials, zero_ptr, zero_words);
  return hipGetLastError();
}

hipError_t
launch_encode(int code_bits, const float* cx, const float* cy, const float* cz, uint64_t n,
              const float* partials, float* bounds, void* codes, uint32_t* hist,
              int digit_bits, int first_pass, int passes, uint32_t* counts, uint32_t groups,
              uint64_t chunk, hipStream_t s)
{
  if (first_pass < 0 || first_pass >= passes)
    return hipErrorInvalidValue;
  const uint32_t words = uint32_t(1u << digit_bits) * uint32_t(passes - first_pass);
  if (words > uint32_t(kHistMaxWords))
    return hipErrorInvalidValue;
  size_t lds = size_t(words) * sizeof(uint32_t);
  uint32_t g = encode_groups(n);
  if (counts) {
    if (groups == 0 || chunk == 0 || uint64_t(groups) * chunk < n)
      return hipErrorInvalidValue;
    g = groups;
  }
  uint32_t* h = hist + size_t(first_pass) * (size_t(1) << digit_bits);
  const uint32_t nparts = centroid_partials_needed(n);
  if (code_bits == 30)
    hipLaunchKernelGGL(k_encode<uint32_t>, dim3(g), dim3(kEncodeBlock), lds, s, cx, cy, cz, n,
                       partials, nparts, bounds, static_cast<uint32_t*>(codes), h, digit_bits,
                       first_pass, passes, counts, chunk);
  else
    hipLaunchKernelGGL(k_encode<uint64_t>, dim3(g), dim3(kEncodeBlock), lds, s, cx, cy, cz, n,
                       partials, nparts, bounds, static_cast<uint64_t*>(codes), h, digit_bits,
                       first_pass, passes, counts, chunk);
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
launch_fit(const float* prims, int kind, uint64_t n, const uint64_t* prim_off, uint32_t nseg,
           void* nodes, const uint32_t* parent, const uint32_t* other_end,
           const uint32_t* leafprim, uint32_t* flags, const void* treelets,
           uint32_t* treelet_count, uint32_t* queue, hipStream_t s)
{
  const Scenes sc = make_scenes(prim_off, nseg, n);
  const uint2* tl = static_cast<const uint2*>(treelets);
  // typical trees hold about 2 n / kTreelet treelets (plus one per small
  // scene); the loops are persistent, so any count is covered
  const uint32_t ga = grid_for(2 * n + uint64_t(sc.nseg) * kTreelet, kTreelet, 768);
  hipLaunchKernelGGL(k_fit_treelets, dim3(ga), dim3(kTreelet), 0, s, prims, kind, sc,
                     static_cast<Node*>(nodes), other_end, leafprim, tl, treelet_count, queue);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  const uint32_t gb = grid_for(4 * n / kTreelet + 64, kBlock, 1024);
  hipLaunchKernelGGL(k_fit_top, dim3(gb), dim3(kBlock), 0, s, prims, kind, sc,
                     static_cast<Node*>(nodes), parent, leafprim, tl, treelet_count, flags);
  return hipGetLastError();
}

// Elements a tree workgroup merges at once before it splits the merge
// (kTreeSlots; LBVH_TREE_GROUP_SLOTS lowers it, so that tests reach the
// split merges with ordinary inputs).
static uint32_t
tree_slot_cap()
{
  const char* v = getenv("LBVH_TREE_GROUP_SLOTS");
  if (!v || !*v)
    return uint32_t(kTreeSlots);
  const long x = strtol(v, nullptr, 10);
  return (x < 1 || x > kTreeSlots) ? uint32_t(kTreeSlots) : uint32_t(x);
}

// Level sizes of k_tree for n positions, and the scratch it needs: the
// segments of every level, then kTreeSubLists split-merge segments per
// workgroup of each level above 0.
static TreeLevels
tree_levels(uint64_t n)
{
  TreeLevels lv{};
  uint32_t nb = uint32_t((n + kTreeLeaves - 1) / kTreeLeaves);
  uint32_t segs = 0, tickets = 0;
  lv.levels = 0;
  for (;;) {
    lv.nblocks[lv.levels] = nb;
    lv.seg_base[lv.levels] = segs;
    lv.ticket_base[lv.levels] = tickets;
    segs += nb;
    lv.levels++;
    if (nb == 1 || lv.levels == 16)
      break;
    const uint32_t groups = (nb + kTreeFan - 1) / kTreeFan;
    tickets += groups;
    nb = groups;
  }
  for (int l = 1; l < lv.levels; l++) {
    lv.tmp_base[l] = segs;
    segs += lv.nblocks[l] * kTreeSubLists;
  }
  lv.tmp_base[0] = segs; // total segments
  lv.slot_cap = uint32_t(kTreeSlots);
  return lv;
}

uint64_t
tree_scratch_bytes(uint64_t n)
{
  const TreeLevels lv = tree_levels(n);
  const uint64_t segs = lv.tmp_base[0];
  // segment counts, then the entries
  const uint64_t head = (segs * 4 + 255) / 256 * 256;
  return head + segs * kTreeSegCap * sizeof(TreeEntry);
}

uint64_t
tree_ticket_words(uint64_t n)
{
  const TreeLevels lv = tree_levels(n);
  return uint64_t(lv.ticket_base[lv.levels - 1]) + 1;
}

template<typename KeyT, typename CodeT, bool SEG>
static void
tree_launch(const void* sorted, const uint32_t* ids, const Scenes& sc, const float* prims,
            int kind, void* nodes, uint32_t* parent, uint32_t* leafprim, uint32_t* flags,
            uint32_t* other_end, void* scratch, uint32_t* tk, hipStream_t s)
{
  const uint64_t n = uint64_t(sc.n);
  TreeLevels lv = tree_levels(n);
  lv.slot_cap = tree_slot_cap();
  const uint64_t segs = lv.tmp_base[0];
  uint32_t* counts = static_cast<uint32_t*>(scratch);
  const uint64_t head = (segs * 4 + 255) / 256 * 256;
  TreeEntry* entries = reinterpret_cast<TreeEntry*>(static_cast<char*>(scratch) + head);
  hipLaunchKernelGGL((k_tree<KeyT, CodeT, SEG, true>), dim3(lv.nblocks[0]), dim3(kTreeLeaves), 0,
                     s, static_cast<const KeyT*>(sorted), ids, sc, prims, kind,
                     static_cast<Node*>(nodes), parent, leafprim, flags, other_end, lv, counts,
                     entries, tk);
  if (lv.levels > 1)
    hipLaunchKernelGGL((k_tree<KeyT, CodeT, SEG, false>), dim3(lv.nblocks[1]), dim3(kTreeUpThreads),
                       0, s, static_cast<const KeyT*>(sorted), ids, sc, prims, kind,
                       static_cast<Node*>(nodes), parent, leafprim, flags, other_end, lv, counts,
                       entries, tk);
}

hipError_t
launch_tree(int code_bits, const void* sorted, const uint32_t* ids, uint64_t n,
            const uint64_t* prim_off, uint32_t nseg, const float* prims, int kind, void* nodes,
            uint32_t* parent, uint32_t* leafprim, uint32_t* flags, uint32_t* other_end,
            void* scratch, uint32_t* tickets, hipStream_t s)
{
  const Scenes sc = make_scenes(prim_off, nseg, n);
  if (tree_levels(n).nblocks[tree_levels(n).levels - 1] != 1)
    return hipErrorInvalidValue; // more levels than the launch descriptor holds
  if (prim_off) {
    if (code_bits != 30)
      return hipErrorInvalidValue;
    tree_launch<uint64_t, uint32_t, true>(sorted, ids, sc, prims, kind, nodes, parent, leafprim,
                                          flags, other_end, scratch, tickets, s);
  } else if (code_bits == 30) {
    tree_launch<uint32_t, uint32_t, false>(sorted, ids, sc, prims, kind, nodes, parent, leafprim,
                                           flags, other_end, scratch, tickets, s);
  } else {
    tree_launch<uint64_t, uint64_t, false>(sorted, ids, sc, prims, kind, nodes, parent, leafprim,
                                           flags, other_end, scratch, tickets, s);
  }
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
