// Closest-hit traversal of builder<double> trees on gfx950.
//
// traverser<double> (lbvh.h:1874-1942) over node<double> records with
// triangle_intersector<double> (lbvh.h:2055-2075), in FP64 throughout:
// make_accel_ray (lbvh.h:991-1010) and the octant box test (lbvh.h:1071-1091,
// hit iff tmax >= max(0, tmin), lbvh.h:1023-1026) on the children's own
// boxes; the root pushed with tmin = +inf; a popped entry skipped when the
// closest hit so far is nearer than its tmin; leaf children intersected at
// once, left then right; hit children pushed far first, near last; a
// 128-entry stack that drops pushes when full (traversal_stack<double,128>,
// lbvh.h:1746-1789).  A leaf value names the triangle<double> record it tests
// (lbvh.h:1904-1905, 1915-1916) and the hit reports that value, u64 with
// UINT64_MAX for a miss (intersection_container<double>, lbvh.h:364-366).
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

// lbvh::node<double> (lbvh.h:220-255): box {min xyz, max xyz}, left, right;
// bit 63 of left / right marks a leaf.
struct alignas(64) Node64
{
  double bmin[3];
  double bmax[3];
  uint64_t left;
  uint64_t right;
};
static_assert(sizeof(Node64) == 64, "node<double> is 64 bytes");

constexpr uint64_t kLeafBit64 = uint64_t(1) << 63;
// Two waves per workgroup: the 1M primary launch 1.10 ms with four, 1.00 ms
// with two (one timed the same).
constexpr int kTraceF64Block = 128;
constexpr int kStackF64 = 128;
// Stack entries kept in LDS per thread (8 KB per 128-thread workgroup).
constexpr int kStackF64Lds = 4;

__device__ __forceinline__ double
dm_min(double a, double b)
{
  return (a < b) ? a : b;
}
__device__ __forceinline__ double
dm_max(double a, double b)
{
  return (a > b) ? a : b;
}

// lbvh::math::dot (lbvh.h:621-630): accumulation starts from 0.
__device__ __forceinline__ double
ddot3(const double* a, const double* b)
{
  double out = 0.0;
  double px = a[0] * b[0];
  out = out + px;
  double py = a[1] * b[1];
  out = out + py;
  double pz = a[2] * b[2];
  out = out + pz;
  return out;
}

// lbvh::math::cross (lbvh.h:609-616)
__device__ __forceinline__ void
dcross3(const double* a, const double* b, double* o)
{
  double ay_bz = a[1] * b[2], az_by = a[2] * b[1];
  double az_bx = a[2] * b[0], ax_bz = a[0] * b[2];
  double ax_by = a[0] * b[1], ay_bx = a[1] * b[0];
  o[0] = ay_bz - az_by;
  o[1] = az_bx - ax_bz;
  o[2] = ax_by - ay_bx;
}

// triangle_intersector<double> on a triangle<double> record {p0, e0, e1,
// normal} (12 doubles) already loaded as six double2.
__device__ __forceinline__ bool
isect_f64(const double2 (&q)[6], const double o[3], const double d[3], double tmin, double tmax,
          double& t_out, double& u_out, double& v_out)
{
  const double2 a = q[0], b = q[1], c = q[2], e = q[3], f = q[4], g = q[5];
  const double p0[3] = { a.x, a.y, b.x };
  const double e0[3] = { b.y, c.x, c.y };
  const double e1[3] = { e.x, e.y, f.x };
  const double nrm[3] = { f.y, g.x, g.y };
  double cc[3] = { p0[0] - o[0], p0[1] - o[1], p0[2] - o[2] };
  double r[3];
  dcross3(d, cc, r);
  double inv_det = 1.0 / ddot3(nrm, d);
  double u = ddot3(r, e1) * inv_det;
  double v = ddot3(r, e0) * inv_det;
  double w = (1.0 - u) - v;
  if ((u >= 0.0) && (v >= 0.0) && (w >= 0.0)) {
    double t = ddot3(nrm, cc) * inv_det;
    if ((t >= tmin) && (t <= tmax)) {
      t_out = t;
      u_out = u;
      v_out = v;
      return true;
    }
  }
  return false;
}

// MODE 0: the first hit found wins ties (the reference's strict <);
// MODE 1: equal distances go to the lowest primitive value.
template<int MODE>
__global__ __launch_bounds__(kTraceF64Block) void
k_trace_f64(const Node64* __restrict__ nodes, const double* __restrict__ tris, uint64_t n,
            const double* __restrict__ rays, uint64_t nrays, uint64_t* __restrict__ prim_out,
            double* __restrict__ t_out, double* __restrict__ uv_out)
{
  const uint64_t i = uint64_t(blockIdx.x) * kTraceF64Block + threadIdx.x;
  if (i >= nrays)
    return;
  const double2* rp = reinterpret_cast<const double2*>(rays + 8 * i);
  const double2 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3];
  const double o[3] = { r0.x, r0.y, r1.x };
  const double d[3] = { r1.y, r2.x, r2.y };
  const double rtmin = r3.x, rtmax = r3.y;
  // make_accel_ray: rcp = 1 / dir, inv_pos = (-pos) * rcp, octant by sign
  double rcp[3], ipos[3];
  bool pos_dir[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    rcp[k] = 1.0 / d[k];
    double np = -o[k];
    ipos[k] = np * rcp[k];
    pos_dir[k] = d[k] > 0.0;
  }
  // The stack: capacity 128 (pushes onto a full stack are dropped).  An
  // entry stands for a pushed node and holds what its own record gave when
  // its box was tested: its two children (packed to 32 bits, bit 31 marking a
  // leaf, as every index is below 2^31) and the box's entry distance (the
  // reference's tmin), so popping it needs no further load.  The first
  // kStackF64Lds entries live in LDS, one column per thread, the rest in a
  // private array; the top entry is kept in registers and reloaded only after
  // a pop empties them (its reload is issued after the children's fetches, so
  // the latencies overlap).
  __shared__ uint2 s_lr[kStackF64Lds * kTraceF64Block];
  __shared__ double s_tmin[kStackF64Lds * kTraceF64Block];
  uint2 spill_lr[kStackF64 - kStackF64Lds];
  double spill_t[kStackF64 - kStackF64Lds];
  auto stk_store = [&](int e, uint2 lr, double tm) {
    if (e < kStackF64Lds) {
      s_lr[e * kTraceF64Block + threadIdx.x] = lr;
      s_tmin[e * kTraceF64Block + threadIdx.x] = tm;
    } else {
      spill_lr[e - kStackF64Lds] = lr;
      spill_t[e - kStackF64Lds] = tm;
    }
  };
  auto stk_load = [&](int e, uint2& lr, double& tm) {
    if (e < kStackF64Lds) {
      lr = s_lr[e * kTraceF64Block + threadIdx.x];
      tm = s_tmin[e * kTraceF64Block + threadIdx.x];
    } else {
      lr = spill_lr[e - kStackF64Lds];
      tm = spill_t[e - kStackF64Lds];
    }
  };
  auto pack = [](uint64_t v) -> uint32_t {
    return (v & kLeafBit64) ? (uint32_t(v) | 0x80000000u) : uint32_t(v);
  };
  uint2 top_lr = make_uint2(0u, 0u);
  double top_t = 0.0;
  int sp = 0;
  bool stale = false; // the registers hold an entry already popped: reload before use
  auto push = [&](uint2 lr, double tm) {
    if (sp >= kStackF64)
      return;
    if (sp > 0 && !stale)
      stk_store(sp - 1, top_lr, top_t);
    stale = false;
    top_lr = lr;
    top_t = tm;
    sp++;
  };
  uint64_t best = ~uint64_t(0);
  double best_t = INFINITY, best_u = 0.0, best_v = 0.0;
  // The root is pushed and popped at once: its record gives the first two
  // children.  A child that would be pushed and popped again straight away
  // (the near one of two, or a single hit child) is taken directly, when its
  // push would have found room, so the visit order and the drop-when-full
  // rule are the reference's.
  uint2 cur = make_uint2(0u, 0u);
  bool have = false;
  // A tree's traversal visits each of its n - 1 internal nodes at most once;
  // a caller's node array that is not a tree (a child cycle) stops loading
  // after n visits and drains its stack, instead of never ending.
  uint32_t visits = 0;
  if (n >= 2) {
    const ulonglong2 lr0 = *reinterpret_cast<const ulonglong2*>(&nodes[0].left);
    cur = make_uint2(pack(lr0.x), pack(lr0.y));
    have = true;
  }
  for (;;) {
    // a pruned entry returns to the loop head by the same path as a node
    // whose children both miss (a second path nests loops in the compiler's
    // structurizer)
    bool visit = true;
    if (!have) {
      if (sp == 0)
        break;
      if (stale)
        stk_load(sp - 1, top_lr, top_t);
      cur = top_lr;
      visit = !(best_t < top_t);
      sp--;
      stale = sp > 0;
    }
    have = false;
    if (!visit || ++visits > n)
      continue;
    // both children's data at once: a leaf's triangle<double> record, an
    // internal node's record; a child out of range (only possible for a
    // corrupt node array) is skipped
    const uint32_t ch[2] = { cur.x, cur.y };
    bool ok[2];
    double2 q[2][6];
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const bool leaf = (ch[side] & 0x80000000u) != 0u;
      const uint64_t idx = ch[side] & 0x7fffffffu;
      ok[side] = leaf ? idx < n : idx + 1 < n;
#pragma unroll
      for (int k = 0; k < 6; k++)
        q[side][k] = make_double2(0.0, 0.0);
      if (ok[side]) {
        const double2* a = reinterpret_cast<const double2*>(leaf ? tris + 12 * idx
                                                                 : &nodes[idx].bmin[0]);
#pragma unroll
        for (int k = 0; k < 4; k++)
          q[side][k] = a[k];
        if (leaf) {
          q[side][4] = a[4];
          q[side][5] = a[5];
        }
      }
    }
    if (stale) { // reload the top while the children are in flight
      stk_load(sp - 1, top_lr, top_t);
      stale = false;
    }
    bool hit2[2] = { false, false };
    double tn2[2] = { 0.0, 0.0 };
    uint2 g2[2] = { make_uint2(0u, 0u), make_uint2(0u, 0u) };
#pragma unroll
    for (int side = 0; side < 2; side++) {
      if (!ok[side])
        continue;
      if (ch[side] & 0x80000000u) {
        const uint64_t idx = ch[side] & 0x7fffffffu;
        double t, u, v;
        if (isect_f64(q[side], o, d, rtmin, rtmax, t, u, v)) {
          const bool take = (MODE == 0) ? (t < best_t)
                                        : ((t < best_t) || (t == best_t && idx < best));
          if (take) {
            best = idx;
            best_t = t;
            best_u = u;
            best_v = v;
          }
        }
      } else {
        const double2 b0 = q[side][0], b1 = q[side][1], b2 = q[side][2], b3 = q[side][3];
        const double lo[3] = { b0.x, b0.y, b1.x };
        const double hi[3] = { b1.y, b2.x, b2.y };
        g2[side] = make_uint2(pack(uint64_t(__double_as_longlong(b3.x))),
                              pack(uint64_t(__double_as_longlong(b3.y))));
        double tn[3], tf[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          double nb = pos_dir[k] ? lo[k] : hi[k];
          double fb = pos_dir[k] ? hi[k] : lo[k];
          double pn = nb * rcp[k];
          tn[k] = pn + ipos[k];
          double pf = fb * rcp[k];
          tf[k] = pf + ipos[k];
        }
        const double tmin = dm_max(dm_max(tn[0], tn[1]), tn[2]);
        const double tmax = dm_min(dm_min(tf[0], tf[1]), tf[2]);
        hit2[side] = tmax >= dm_max(0.0, tmin);
        tn2[side] = tmin;
      }
    }
    uint2 next;
    double next_t;
    if (hit2[0] && hit2[1]) {
      const bool left_near = tn2[0] < tn2[1];
      push(left_near ? g2[1] : g2[0], left_near ? tn2[1] : tn2[0]);
      next = left_near ? g2[0] : g2[1];
      next_t = left_near ? tn2[0] : tn2[1];
    } else if (hit2[0] || hit2[1]) {
      next = hit2[0] ? g2[0] : g2[1];
      next_t = hit2[0] ? tn2[0] : tn2[1];
    } else {
      continue;
    }
    // push-then-pop of the near child: its push needs room; its pop is then
    // pruned when a closer hit is already known
    if (sp < kStackF64 && !(best_t < next_t)) {
      cur = next;
      have = true;
    }
  }
  const bool hit = best != ~uint64_t(0);
  prim_out[i] = best;
  t_out[i] = best_t;
  if (uv_out)
    *reinterpret_cast<double2*>(uv_out + 2 * i) = make_double2(hit ? best_u : 0.0,
                                                               hit ? best_v : 0.0);
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

hipError_t
launch_trace_f64(int mode, const void* nodes, const double* tris, uint64_t n,
                 const double* rays, uint64_t nrays, uint64_t* prim, double* t, double* uv,
                 hipStream_t s)
{
  if (nrays == 0)
    return hipSuccess;
  const uint32_t g = uint32_t((nrays + kTraceF64Block - 1) / kTraceF64Block);
  const Node64* nd = static_cast<const Node64*>(nodes);
  if (mode == 0)
    hipLaunchKernelGGL((k_trace_f64<0>), dim3(g), dim3(kTraceF64Block), 0, s, nd, tris, n, rays,
                       nrays, prim, t, uv);
  else if (mode == 1)
    hipLaunchKernelGGL((k_trace_f64<1>), dim3(g), dim3(kTraceF64Block), 0, s, nd, tris, n, rays,
                       nrays, prim, t, uv);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

} // namespace lbvh_eng
