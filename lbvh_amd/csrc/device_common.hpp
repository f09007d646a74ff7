// Device-side building blocks shared by the gfx950 kernels.
//
// Arithmetic here reproduces the reference CPU path bit for bit, so every
// helper states the reference expression it mirrors.  The library is built
// with -ffp-contract=off (no FMA fusion), IEEE division and square root, and
// denormals preserved.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbvh_dev {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kWave = 64;

// lbvh::node<float> (lbvh.h:220-255): box {min xyz, max xyz}, left, right.
struct alignas(32) Node
{
  float bmin[3];
  float bmax[3];
  uint32_t left;
  uint32_t right;
};
static_assert(sizeof(Node) == 32, "node<float> is 32 bytes");

struct Box
{
  float lo[3];
  float hi[3];
};

// libstdc++ std::min(a, b) = (b < a) ? b : a — used by the vector min/max of
// lbvh::math (lbvh.h:638-660) and hence by union_of and the AABB converter.
__device__ __forceinline__ float
std_min(float a, float b)
{
  return (b < a) ? b : a;
}
// std::max(a, b) = (a < b) ? b : a
__device__ __forceinline__ float
std_max(float a, float b)
{
  return (a < b) ? b : a;
}
// lbvh::math::min / max (lbvh.h:567-580): ternaries, used by the box test.
__device__ __forceinline__ float
m_min(float a, float b)
{
  return (a < b) ? a : b;
}
__device__ __forceinline__ float
m_max(float a, float b)
{
  return (a > b) ? a : b;
}

// triangle_aabb_converter (lbvh.h:2001-2012): p1 = p0 - e0, p2 = p0 + e1,
// min(min(p0, p1), p2), max(max(p0, p1), p2) per component.
__device__ __forceinline__ Box
tri_box(const float p0[3], const float e0[3], const float e1[3])
{
  Box b;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float p1 = p0[k] - e0[k];
    float p2 = p0[k] + e1[k];
    float mn = std_min(p0[k], p1);
    float mx = std_max(p0[k], p1);
    b.lo[k] = std_min(mn, p2);
    b.hi[k] = std_max(mx, p2);
  }
  return b;
}

// union_of (lbvh.h:1208-1213): left operand first.
__device__ __forceinline__ Box
box_union(const Box& a, const Box& b)
{
  Box o;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    o.lo[k] = std_min(a.lo[k], b.lo[k]);
    o.hi[k] = std_max(a.hi[k], b.hi[k]);
  }
  return o;
}

// Primitive input kinds: 48-byte triangles or 24-byte precomputed boxes.
enum PrimKind : int
{
  kPrimTriangle = 0,
  kPrimAabb = 1
};

__device__ __forceinline__ Box
load_prim_box(const float* __restrict__ prims, int kind, uint64_t i)
{
  if (kind == kPrimTriangle) {
    const float4* p = reinterpret_cast<const float4*>(prims + 12 * i);
    float4 a = p[0]; // p0.xyz, e0.x
    float4 b = p[1]; // e0.yz, e1.xy
    float2 c = *reinterpret_cast<const float2*>(prims + 12 * i + 8); // e1.z, n.x
    float p0[3] = { a.x, a.y, a.z };
    float e0[3] = { a.w, b.x, b.y };
    float e1[3] = { b.z, b.w, c.x };
    return tri_box(p0, e0, e1);
  }
  const float2* p = reinterpret_cast<const float2*>(prims + 6 * i);
  float2 a = p[0], b = p[1], c = p[2];
  Box o;
  o.lo[0] = a.x;
  o.lo[1] = a.y;
  o.lo[2] = b.x;
  o.hi[0] = b.y;
  o.hi[1] = c.x;
  o.hi[2] = c.y;
  return o;
}

// ---- scene batches ---------------------------------------------------------
// A batch concatenates scenes: scene s owns primitives (and sorted positions)
// [P, P + n_s) with P = prim_off[s], internal nodes [P - s, P - s + n_s - 1)
// and parent links [2P - s, 2P - s + 2 n_s - 1); every n_s >= 2.  A single
// scene is prim_off == nullptr, n_0 = n.
struct Scenes
{
  const uint64_t* prim_off; // [nseg + 1] or nullptr
  uint32_t nseg;
  int64_t n; // all primitives
};

// Largest s in [0, nseg) with key(s) <= x, where key(s) = off[s] * mul -
// (minus_s ? s : 0) is strictly increasing and key(0) = 0: the scene of a
// primitive position (mul 1), an internal node (mul 1, minus_s), a parent
// link (mul 2, minus_s), or of a ray given per-scene ray offsets (mul 1).
__device__ __forceinline__ uint32_t
find_scene(const uint64_t* __restrict__ off, uint32_t nseg, int64_t x, int mul, bool minus_s)
{
  uint32_t lo = 0, hi = nseg; // key(lo) <= x < key(hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    const int64_t key = int64_t(off[mid]) * mul - (minus_s ? int64_t(mid) : 0);
    if (key <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Where one scene's arrays start inside the batch.
struct SceneView
{
  int64_t P;  // first primitive / sorted position
  int64_t n;  // primitives
  int64_t nb; // first internal node
  int64_t lb; // first parent link
};

__device__ __forceinline__ SceneView
scene_view(const Scenes& sc, uint32_t seg)
{
  SceneView v;
  if (sc.prim_off) {
    v.P = int64_t(sc.prim_off[seg]);
    v.n = int64_t(sc.prim_off[seg + 1]) - v.P;
  } else {
    seg = 0;
    v.P = 0;
    v.n = sc.n;
  }
  v.nb = v.P - int64_t(seg);
  v.lb = 2 * v.P - int64_t(seg);
  return v;
}

__device__ __forceinline__ int
prim_stride(int kind)
{
  return kind == kPrimTriangle ? 12 : 6;
}

// ---- inter-workgroup communication ----------------------------------------
// Stores and loads that must be seen by workgroups on other XCDs use relaxed
// agent-scope atomics (device-coherent, bypassing the per-XCD L2 for the line
// state); a vector-memory wait orders them before a following flag atomic.
// This avoids the whole-L2 write-back/invalidate that an agent-scope fence
// costs on gfx950.

__device__ __forceinline__ void
coherent_store_u32(uint32_t* p, uint32_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t
coherent_load_u32(const uint32_t* p)
{
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void
coherent_store_u64(uint64_t* p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t
coherent_load_u64(const uint64_t* p)
{
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void
coherent_store_f32(float* p, float v)
{
  coherent_store_u32(reinterpret_cast<uint32_t*>(p), __float_as_uint(v));
}
__device__ __forceinline__ float
coherent_load_f32(const float* p)
{
  return __uint_as_float(coherent_load_u32(reinterpret_cast<const uint32_t*>(p)));
}
// Waits until this wave's outstanding vector memory operations (including
// the coherent stores above) have completed; also a compiler barrier.
__device__ __forceinline__ void
vmem_drain()
{
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}
__device__ __forceinline__ uint32_t
atomic_add_agent(uint32_t* p, uint32_t v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- wave helpers ----------------------------------------------------------

__device__ __forceinline__ uint64_t
ballot(bool p)
{
  return __ballot(p);
}
__device__ __forceinline__ uint64_t
lanemask_lt()
{
  uint32_t lane = __lane_id();
  return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}
__device__ __forceinline__ int
popc64(uint64_t x)
{
  return __popcll(x);
}

} // namespace lbvh_dev
