// Closest-hit traversal on gfx950.
//
// Reproduces traverser::operator() (lbvh.h:1874-1942) exactly: the root is
// pushed with tmin = +inf; a popped entry is skipped when the closest hit so
// far is nearer than its tmin; leaf children are intersected at once, left
// then right; internal children get the octant slab test (make_accel_ray,
// lbvh.h:991-1010; detail::intersect, lbvh.h:1071-1091; hit iff
// tmax >= max(0, tmin), lbvh.h:1023-1026) and are pushed far first, near
// last; the stack holds 128 entries and drops pushes when full
// (traversal_stack<float,128>, lbvh.h:1746-1789).
//
// It walks the reference's node array itself.  A node is visited with its
// two children already known (they came with its own record, loaded when its
// box was tested), so one round of independent loads serves a visit: the
// 32-byte record of an internal child (its box and its children) and the
// 48-byte triangle of a leaf child are fetched together.  A pushed node
// carries its children on the stack.
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

using lbvh_eng::TraceJob;

// Two waves per workgroup: measured faster than four (the clustered 1M
// primary launch 0.68 -> 0.65 ms) and the same as one.
constexpr int kTraceBlock = 128;
// The occupancy knob's "no limit" (the registers then allow what they allow).
constexpr int kTraceMaxWavesPerSimd = 8;
constexpr int kStack = 128;
// Stack entries below this depth live in LDS (12 bytes each: 9 KB per
// workgroup, so the registers, not LDS, limit the waves per CU); deeper ones
// in the private array.
constexpr int kStackLds = 6;

struct Ray
{
  float o[3];
  float d[3];
  float tmin, tmax;
};

// lbvh::math::dot (lbvh.h:621-630): accumulation starts from 0.
__device__ __forceinline__ float
dot3(const float* a, const float* b)
{
  float out = 0.0f;
  float px = a[0] * b[0];
  out = out + px;
  float py = a[1] * b[1];
  out = out + py;
  float pz = a[2] * b[2];
  out = out + pz;
  return out;
}

// lbvh::math::cross (lbvh.h:609-616)
__device__ __forceinline__ void
cross3(const float* a, const float* b, float* o)
{
  float ay_bz = a[1] * b[2], az_by = a[2] * b[1];
  float az_bx = a[2] * b[0], ax_bz = a[0] * b[2];
  float ax_by = a[0] * b[1], ay_bx = a[1] * b[0];
  o[0] = ay_bz - az_by;
  o[1] = az_bx - ax_bz;
  o[2] = ax_by - ay_bx;
}

struct Tri
{
  float p0[3], e0[3], e1[3], nrm[3];
};

// triangle_intersector (lbvh.h:2055-2075): two-sided, not watertight.
__device__ __forceinline__ bool
isect_reference(const Tri& tr, const Ray& r, float& t_out, float& u_out, float& v_out)
{
  float c[3] = { tr.p0[0] - r.o[0], tr.p0[1] - r.o[1], tr.p0[2] - r.o[2] };
  float rr[3];
  cross3(r.d, c, rr);
  float inv_det = 1.0f / dot3(tr.nrm, r.d);
  float u = dot3(rr, tr.e1) * inv_det;
  float v = dot3(rr, tr.e0) * inv_det;
  float w = (1.0f - u) - v;
  if ((u >= 0.0f) && (v >= 0.0f) && (w >= 0.0f)) {
    float t = dot3(tr.nrm, c) * inv_det;
    if ((t >= r.tmin) && (t <= r.tmax)) {
      t_out = t;
      u_out = u;
      v_out = v;
      return true;
    }
  }
  return false;
}

// Watertight test (Woop, Benthin, Wald, JCGT 2(1) 2013) on A = p0,
// B = p0 - e0, C = p0 + e1; u, v are the weights of B and C.
__device__ __forceinline__ bool
isect_watertight(const Tri& tr, const Ray& r, float& t_out, float& u_out, float& v_out)
{
  float ax = fabsf(r.d[0]), ay = fabsf(r.d[1]), az = fabsf(r.d[2]);
  int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
  int kx = (kz == 2) ? 0 : kz + 1;
  int ky = (kx == 2) ? 0 : kx + 1;
  float dkz = r.d[0], dkx = r.d[0], dky = r.d[0];
  // component selects without dynamic register indexing
  auto comp = [](const float* v, int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); };
  dkz = comp(r.d, kz);
  if (dkz < 0.0f) {
    int tmp = kx;
    kx = ky;
    ky = tmp;
  }
  dkx = comp(r.d, kx);
  dky = comp(r.d, ky);
  float Sx = dkx / dkz;
  float Sy = dky / dkz;
  float Sz = 1.0f / dkz;
  float va[3], vb[3], vc[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float b = tr.p0[k] - tr.e0[k];
    float c = tr.p0[k] + tr.e1[k];
    va[k] = tr.p0[k] - r.o[k];
    vb[k] = b - r.o[k];
    vc[k] = c - r.o[k];
  }
  float az_ = comp(va, kz), bz_ = comp(vb, kz), cz_ = comp(vc, kz);
  float sxa = Sx * az_, sya = Sy * az_;
  float sxb = Sx * bz_, syb = Sy * bz_;
  float sxc = Sx * cz_, syc = Sy * cz_;
  float Ax = comp(va, kx) - sxa, Ay = comp(va, ky) - sya;
  float Bx = comp(vb, kx) - sxb, By = comp(vb, ky) - syb;
  float Cx = comp(vc, kx) - sxc, Cy = comp(vc, ky) - syc;
  float cb = Cx * By, cyb = Cy * Bx;
  float U = cb - cyb;
  float acy = Ax * Cy, ayc = Ay * Cx;
  float V = acy - ayc;
  float bay = Bx * Ay, bya = By * Ax;
  float W = bay - bya;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
    return false;
  float det = U + V;
  det = det + W;
  if (det == 0.0f)
    return false;
  float Az = Sz * az_, Bz = Sz * bz_, Cz = Sz * cz_;
  float ua = U * Az, vb2 = V * Bz, wc = W * Cz;
  float T = ua + vb2;
  T = T + wc;
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t >= r.tmin && t <= r.tmax))
    return false;
  t_out = t;
  u_out = V * rcp;
  v_out = W * rcp;
  return true;
}

template<int MODE>
__device__ __forceinline__ bool
better(float t, uint32_t id, float best_t, uint32_t best_id)
{
  if (MODE == 0)
    return t < best_t;
  return (t < best_t) || (t == best_t && id < best_id && t != INFINITY);
}

// Ray layouts: kLayoutSingle, one tree; kLayoutBatch, a scene batch (ray i
// belongs to the scene s with ray_off[s] <= i < ray_off[s + 1], whose nodes
// start at prim_off[s] - s and whose triangles start at prim_off[s]; ids are
// scene-local); kLayoutJobs, independent trees with their own ray and output
// buffers (job k covers global rays [jobs[k].ray_begin, jobs[k + 1].ray_begin)).
constexpr int kLayoutSingle = 0;
constexpr int kLayoutBatch = 1;
constexpr int kLayoutJobs = 2;

// The last job with ray_begin <= i.
__device__ __forceinline__ uint32_t
job_of(const TraceJob* __restrict__ jobs, uint32_t njobs, uint64_t i)
{
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (jobs[mid].ray_begin <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// order (single layout only, may be null): ray k of the launch is ray
// order[k] of the input, and its hit is written there; a permutation that
// puts rays which visit the same nodes into the same wave.
template<int MODE, int LAYOUT>
__global__ __launch_bounds__(kTraceBlock) void
k_trace(const Node* __restrict__ nodes, const float* __restrict__ tris, uint32_t nprims,
        const uint64_t* __restrict__ prim_off, const uint64_t* __restrict__ ray_off,
        uint32_t nseg, const TraceJob* __restrict__ jobs, const float* __restrict__ rays,
        const uint32_t* __restrict__ order, uint64_t nrays, uint32_t* __restrict__ prim_out,
        float* __restrict__ t_out, float* __restrict__ uv_out)
{
  // batches and jobs: the ray offsets of the scenes / jobs this workgroup's
  // rays fall in, staged once (s_rb) when they are at most kTraceBlock
  constexpr bool kStage = LAYOUT != kLayoutSingle;
  __shared__ uint64_t s_rb[kStage ? kTraceBlock + 1 : 1];
  __shared__ uint32_t s_first, s_span;
  const uint64_t i0 = uint64_t(blockIdx.x) * kTraceBlock;
  uint64_t i = i0 + threadIdx.x;
  uint32_t first = 0, span = 0;
  if constexpr (kStage) {
    if (threadIdx.x == 0) {
      const uint64_t last = (i0 + kTraceBlock <= nrays ? i0 + kTraceBlock : nrays) - 1;
      uint32_t a, b;
      if constexpr (LAYOUT == kLayoutBatch) {
        a = find_scene(ray_off, nseg, int64_t(i0), 1, false);
        b = find_scene(ray_off, nseg, int64_t(last), 1, false);
      } else {
        a = job_of(jobs, nseg, i0);
        b = job_of(jobs, nseg, last);
      }
      s_first = a;
      s_span = (b - a + 1 <= uint32_t(kTraceBlock)) ? b - a + 1 : 0u; // 0: not staged
    }
    __syncthreads();
    first = s_first;
    span = s_span;
    for (uint32_t t = threadIdx.x; t < span; t += kTraceBlock)
      s_rb[t] = (LAYOUT == kLayoutBatch) ? ray_off[first + t] : jobs[first + t].ray_begin;
    __syncthreads();
  }
  if (i >= nrays)
    return;
  uint32_t sel = 0; // the ray's scene or job
  if constexpr (kStage) {
    if (span) {
      uint32_t lo = 0, hi = span; // s_rb[lo] <= i, the last such entry
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_rb[mid] <= i)
          lo = mid;
        else
          hi = mid;
      }
      sel = first + lo;
    } else {
      sel = (LAYOUT == kLayoutBatch) ? find_scene(ray_off, nseg, int64_t(i), 1, false)
                                     : job_of(jobs, nseg, i);
    }
  }
  if constexpr (LAYOUT == kLayoutBatch) {
    const uint32_t sc = sel;
    const uint64_t P = prim_off[sc];
    nprims = uint32_t(prim_off[sc + 1] - P);
    nodes += P - sc;
    tris += 12 * P;
  } else if constexpr (LAYOUT == kLayoutJobs) {
    const TraceJob& jb = jobs[sel];
    nodes = static_cast<const Node*>(jb.nodes);
    tris = jb.tris;
    nprims = jb.nprims;
    rays = jb.rays;
    prim_out = jb.prim;
    t_out = jb.t;
    uv_out = jb.uv;
    i -= jb.ray_begin;
  } else {
    if (order)
      i = order[i];
  }
  const float4* rp = reinterpret_cast<const float4*>(rays + 8 * i);
  const float4 r0 = rp[0], r1 = rp[1];
  Ray r;
  r.o[0] = r0.x;
  r.o[1] = r0.y;
  r.o[2] = r0.z;
  r.d[0] = r0.w;
  r.d[1] = r1.x;
  r.d[2] = r1.y;
  r.tmin = r1.z;
  r.tmax = r1.w;

  // make_accel_ray
  float rcp[3], ipos[3];
  bool pos_dir[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    rcp[k] = 1.0f / r.d[k];
    float np = -r.o[k];
    ipos[k] = np * rcp[k];
    pos_dir[k] = r.d[k] > 0.0f;
  }

  // The traversal stack: capacity 128 with pushes dropped when full, as the
  // reference's.  An entry stands for a pushed node and holds what its own
  // node record gave when its box was tested: its two children and the box's
  // entry distance (the reference's tmin), so popping it needs no further
  // load.  Entries 0 .. sp - 2 are in LDS (the first kStackLds, one column
  // per thread and field) and a private array (the rest, the only thing in
  // private memory); the top one is kept in registers: a pop hands it over at
  // once and leaves the register stale (stale: the top is still entry
  // sp - 1).  The reload is issued after the children's fetches, so the
  // latencies overlap; a push onto a stale top needs no reload at all.
  __shared__ uint32_t s_stk[3 * kStackLds * kTraceBlock];
  uint3 spill[kStack - kStackLds];
  auto stk_store = [&](uint32_t e, uint3 v) {
    if (e < uint32_t(kStackLds)) {
      uint32_t* q = &s_stk[e * kTraceBlock + threadIdx.x];
      q[0] = v.x;
      q[kStackLds * kTraceBlock] = v.y;
      q[2 * kStackLds * kTraceBlock] = v.z;
    } else {
      spill[e - kStackLds] = v;
    }
  };
  auto stk_load = [&](uint32_t e) -> uint3 {
    if (e < uint32_t(kStackLds)) {
      const uint32_t* q = &s_stk[e * kTraceBlock + threadIdx.x];
      return make_uint3(q[0], q[kStackLds * kTraceBlock], q[2 * kStackLds * kTraceBlock]);
    }
    return spill[e - kStackLds];
  };
  uint3 top = make_uint3(0u, 0u, 0u);
  uint32_t sp = 0;
  bool stale = false;
  auto push = [&](uint32_t l, uint32_t rr, float tn) {
    if (sp >= kStack)
      return;
    if (sp > 0 && !stale)
      stk_store(sp - 1, top);
    stale = false;
    top = make_uint3(l, rr, __float_as_uint(tn));
    sp++;
  };
  auto pop = [&]() {
    if (stale)
      top = stk_load(sp - 1);
    const uint3 e = top;
    sp--;
    stale = sp > 0;
    return e;
  };

  uint32_t best = 0xffffffffu;
  float best_t = INFINITY, best_u = 0.0f, best_v = 0.0f;

  // The reference pushes the root and pops it at once: its record gives the
  // first two children.  A node it would push and pop again straight away
  // (the near child, or a single hit child) is taken directly here, but only
  // when that push would have found room, so the visit order and the
  // drop-when-full rule are unchanged.
  const float* nodes_f = reinterpret_cast<const float*>(nodes);
  uint32_t cl = 0u, cr = 0u;
  bool have = false;
  if (nprims >= 2u) {
    const float4 q1 = reinterpret_cast<const float4*>(nodes_f)[1];
    cl = __float_as_uint(q1.z);
    cr = __float_as_uint(q1.w);
    have = true;
  }
  for (;;) {
    // A pruned entry goes back to the loop head by the same path as a node
    // whose children both miss: with a separate path, the compiler's
    // structurizer nests loops and duplicates the child tests (measured 12%
    // slower on the clustered 1M scene).
    bool visit = true;
    if (!have) {
      if (sp == 0)
        break;
      const uint3 e = pop();
      visit = !(best_t < __uint_as_float(e.z));
      cl = e.x;
      cr = e.y;
    }
    have = false;
    if (!visit)
      continue;
    // Both children's data at once: a leaf's triangle, an internal node's
    // record (its box and children).  A child out of range (only possible
    // for a corrupt node array) is skipped.
    const uint32_t ch[2] = { cl, cr };
    bool ok[2];
    float4 q[2][3];
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const uint32_t c = ch[side];
      const bool leaf = (c & kLeafBit) != 0u;
      ok[side] = leaf ? ((c & ~kLeafBit) < nprims) : (c + 1 < nprims);
      q[side][0] = q[side][1] = q[side][2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (ok[side]) {
        const float4* a = reinterpret_cast<const float4*>(
          leaf ? tris + 12 * uint64_t(c & ~kLeafBit) : nodes_f + 8 * uint64_t(c));
        q[side][0] = a[0];
        q[side][1] = a[1];
        if (leaf)
          q[side][2] = a[2];
      }
    }
    if (stale) { // reload the top while the children are in flight
      top = stk_load(sp - 1);
      stale = false;
    }
    bool hit2[2] = { false, false };
    float tn2[2] = { 0.0f, 0.0f };
    uint32_t gl[2] = { 0u, 0u }, gr[2] = { 0u, 0u };
#pragma unroll
    for (int side = 0; side < 2; side++) {
      if (!ok[side])
        continue;
      const float4 a = q[side][0], b = q[side][1], c = q[side][2];
      if (ch[side] & kLeafBit) {
        const uint32_t p = ch[side] & ~kLeafBit;
        Tri tr;
        tr.p0[0] = a.x;
        tr.p0[1] = a.y;
        tr.p0[2] = a.z;
        tr.e0[0] = a.w;
        tr.e0[1] = b.x;
        tr.e0[2] = b.y;
        tr.e1[0] = b.z;
        tr.e1[1] = b.w;
        tr.e1[2] = c.x;
        tr.nrm[0] = c.y;
        tr.nrm[1] = c.z;
        tr.nrm[2] = c.w;
        float t, u, v;
        bool hit = (MODE == 2) ? isect_watertight(tr, r, t, u, v) : isect_reference(tr, r, t, u, v);
        if (hit && better<MODE>(t, p, best_t, best)) {
          best = p;
          best_t = t;
          best_u = u;
          best_v = v;
        }
      } else {
        const float lo[3] = { a.x, a.y, a.z };
        const float hi[3] = { a.w, b.x, b.y };
        gl[side] = __float_as_uint(b.z);
        gr[side] = __float_as_uint(b.w);
        float tn[3], tf[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          float nb = pos_dir[k] ? lo[k] : hi[k];
          float fb = pos_dir[k] ? hi[k] : lo[k];
          float pn = nb * rcp[k];
          tn[k] = pn + ipos[k];
          float pf = fb * rcp[k];
          tf[k] = pf + ipos[k];
        }
        float tmin = m_max(m_max(tn[0], tn[1]), tn[2]);
        float tmax = m_min(m_min(tf[0], tf[1]), tf[2]);
        hit2[side] = tmax >= m_max(0.0f, tmin);
        tn2[side] = tmin;
      }
    }
    const bool lh = hit2[0], rh = hit2[1];
    uint32_t next_l, next_r;
    float next_t;
    if (lh && rh) {
      const bool left_near = tn2[0] < tn2[1];
      const int far = left_near ? 1 : 0;
      push(gl[far], gr[far], tn2[far]);
      next_l = gl[1 - far];
      next_r = gr[1 - far];
      next_t = tn2[1 - far];
    } else if (lh || rh) {
      next_l = lh ? gl[0] : gl[1];
      next_r = lh ? gr[0] : gr[1];
      next_t = lh ? tn2[0] : tn2[1];
    } else {
      continue;
    }
    // push-then-pop of the near child: its push needs room; its pop is then
    // pruned when a closer hit is already known
    if (sp < kStack && !(best_t < next_t)) {
      cl = next_l;
      cr = next_r;
      have = true;
    }
  }
  prim_out[i] = best;
  t_out[i] = best_t;
  if (uv_out)
    *reinterpret_cast<float2*>(uv_out + 2 * i) = make_float2(best_u, best_v);
}

// ---- traversal order of a ray batch ---------------------------------------------
// Key of a ray: its direction octant (3 bits) above the 12-bit Morton code of
// the point o + d * reach, quantized to 16 cells per axis over the tree's
// centroid bounds grown by reach on every side, reach being the bounds'
// largest extent.  Rays with one key start in the same part of the scene and
// head the same way (for primary rays from one eye: neighbouring pixels), so
// a wave of them visits much the same nodes in the same order.  Sorted with
// two 8-bit radix passes; each workgroup encodes one chunk of the
// reduce-then-scan sort and writes that chunk's first-pass digit counts and
// its share of both passes' histograms (zero before).

constexpr int kRayKeyBlock = 256;

__device__ __forceinline__ uint32_t
spread4(uint32_t v)
{
  return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6);
}

__global__ __launch_bounds__(kRayKeyBlock) void
k_ray_keys(const float* __restrict__ rays, uint64_t nrays, const float* __restrict__ bounds,
           uint32_t* __restrict__ keys, uint32_t* __restrict__ hist,
           uint32_t* __restrict__ counts, uint64_t chunk)
{
  __shared__ uint32_t s_hist[2 * 256];
  for (int j = threadIdx.x; j < 2 * 256; j += kRayKeyBlock)
    s_hist[j] = 0u;
  float lo[3], scale[3];
  float reach = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; k++)
    reach = fmaxf(reach, bounds[3 + k] - bounds[k]);
  reach = fmaxf(reach, 1e-20f);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = bounds[k] - reach;
    scale[k] = 16.0f / (bounds[3 + k] + reach - lo[k]);
  }
  __syncthreads();
  const uint64_t b = uint64_t(blockIdx.x) * chunk;
  const uint64_t end = (b + chunk < nrays) ? b + chunk : nrays;
  for (uint64_t i = b + threadIdx.x; i < end; i += kRayKeyBlock) {
    const float4* rp = reinterpret_cast<const float4*>(rays + 8 * i);
    const float4 r0 = rp[0], r1 = rp[1];
    const float o[3] = { r0.x, r0.y, r0.z };
    const float d[3] = { r0.w, r1.x, r1.y };
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float p = o[k] + d[k] * reach;
      const float c = fminf(fmaxf((p - lo[k]) * scale[k], 0.0f), 15.0f); // NaN -> 0
      q[k] = uint32_t(c);
    }
    const uint32_t oct = (d[0] < 0.0f ? 4u : 0u) | (d[1] < 0.0f ? 2u : 0u) | (d[2] < 0.0f ? 1u : 0u);
    const uint32_t key = (oct << 12) | (spread4(q[0]) << 2) | (spread4(q[1]) << 1) | spread4(q[2]);
    keys[i] = key;
    atomicAdd(&s_hist[key & 0xffu], 1u);
    atomicAdd(&s_hist[256 + (key >> 8)], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 2 * 256; j += kRayKeyBlock) {
    const uint32_t v = s_hist[j];
    if (v)
      atomicAdd(&hist[j], v);
  }
  for (int dd = threadIdx.x; dd < 256; dd += kRayKeyBlock)
    counts[size_t(dd) * gridDim.x + blockIdx.x] = s_hist[dd];
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

// LDS a workgroup reserves beyond its own static `own` bytes so that at most
// `waves` waves share a SIMD: a workgroup puts its kTraceBlock / 64 waves on
// as many of the CU's 4 SIMDs, so that is waves * 4 / (kTraceBlock / 64)
// workgroups per CU, and one more must not fit.  0 = no limit beyond the
// registers'.
static size_t
occupancy_lds(int waves, size_t own)
{
  if (waves <= 0 || waves >= kTraceMaxWavesPerSimd)
    return 0;
  constexpr size_t kLdsPerCu = 160 * 1024;
  const size_t wgs = size_t(waves) * 4 / size_t(kTraceBlock / kWave);
  const size_t per_wg = (kLdsPerCu / wgs) & ~size_t(511);
  return per_wg > own ? per_wg - own : 0;
}

// Static LDS of a kernel, from its attributes (0 if they cannot be read; the
// failed query's error is cleared, so that it is not reported by the launch).
template<typename K>
static size_t
static_lds(K kernel)
{
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kernel)) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return a.sharedSizeBytes;
}

template<int MODE>
static void
trace_mode(uint32_t g, int waves, hipStream_t s, const Node* nd, const float* tris,
           uint64_t n, const uint64_t* prim_off, const uint64_t* ray_off, uint32_t nseg,
           const TraceJob* jobs, const float* rays, const uint32_t* order, uint64_t nrays,
           uint32_t* prim, float* t, float* uv)
{
  static const size_t own_jobs = static_lds(k_trace<MODE, kLayoutJobs>);
  static const size_t own_batch = static_lds(k_trace<MODE, kLayoutBatch>);
  static const size_t own_single = static_lds(k_trace<MODE, kLayoutSingle>);
  if (jobs)
    hipLaunchKernelGGL((k_trace<MODE, kLayoutJobs>), dim3(g), dim3(kTraceBlock),
                       occupancy_lds(waves, own_jobs), s, nullptr, nullptr, 0u, nullptr, nullptr,
                       nseg, jobs, nullptr, nullptr, nrays, nullptr, nullptr, nullptr);
  else if (prim_off)
    hipLaunchKernelGGL((k_trace<MODE, kLayoutBatch>), dim3(g), dim3(kTraceBlock),
                       occupancy_lds(waves, own_batch), s, nd, tris, uint32_t(n), prim_off,
                       ray_off, nseg, nullptr, rays, nullptr, nrays, prim, t, uv);
  else
    hipLaunchKernelGGL((k_trace<MODE, kLayoutSingle>), dim3(g), dim3(kTraceBlock),
                       occupancy_lds(waves, own_single), s, nd, tris, uint32_t(n), nullptr,
                       nullptr, 1u, nullptr, rays, order, nrays, prim, t, uv);
}

static hipError_t
trace_dispatch(int mode, const void* nodes, const float* tris, uint64_t n, const uint64_t* prim_off,
               const uint64_t* ray_off, uint32_t nseg, const TraceJob* jobs, const float* rays,
               const uint32_t* order, uint64_t nrays, uint32_t* prim, float* t, float* uv,
               int waves, hipStream_t s)
{
  if (nrays == 0)
    return hipSuccess;
  const uint32_t g = uint32_t((nrays + kTraceBlock - 1) / kTraceBlock);
  const Node* nd = static_cast<const Node*>(nodes);
#define LBVH_TM(M)                                                                                \
  trace_mode<M>(g, waves, s, nd, tris, n, prim_off, ray_off, nseg, jobs, rays, order, nrays,       \
                prim, t, uv)
  switch (mode) {
    case 0:
      LBVH_TM(0);
      break;
    case 1:
      LBVH_TM(1);
      break;
    case 2:
      LBVH_TM(2);
      break;
    default:
      return hipErrorInvalidValue;
  }
#undef LBVH_TM
  return hipGetLastError();
}

hipError_t
launch_trace(int mode, const void* nodes, const float* tris, uint64_t n, const uint64_t* prim_off,
             const uint64_t* ray_off, uint32_t nseg, const float* rays, uint64_t nrays,
             uint32_t* prim, float* t, float* uv, hipStream_t s, const uint32_t* order,
             int waves_per_simd)
{
  if (prim_off && (!ray_off || nseg == 0))
    return hipErrorInvalidValue;
  if (prim_off && order)
    return hipErrorInvalidValue;
  return trace_dispatch(mode, nodes, tris, n, prim_off, ray_off, nseg, nullptr, rays, order, nrays,
                        prim, t, uv, waves_per_simd, s);
}

hipError_t
launch_trace_jobs(int mode, const TraceJob* jobs, uint32_t njobs, uint64_t nrays, hipStream_t s)
{
  if (!jobs || njobs == 0)
    return hipErrorInvalidValue;
  return trace_dispatch(mode, nullptr, nullptr, 0, nullptr, nullptr, njobs, jobs, nullptr, nullptr,
                        nrays, nullptr, nullptr, nullptr, 0, s);
}

hipError_t
launch_ray_keys(const float* rays, uint64_t nrays, const float* bounds, uint32_t* keys,
                uint32_t* hist, uint32_t* counts, uint32_t groups, uint64_t chunk, hipStream_t s)
{
  if (nrays == 0 || groups == 0 || chunk == 0 || uint64_t(groups) * chunk < nrays)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ray_keys, dim3(groups), dim3(kRayKeyBlock), 0, s, rays, nrays, bounds,
                     keys, hist, counts, chunk);
  return hipGetLastError();
}

} // namespace lbvh_eng
