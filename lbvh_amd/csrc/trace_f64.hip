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
constexpr int kTraceF64Block = 256;
constexpr int kStackF64 = 128;
// Stack entries kept in LDS per thread (12 KB per 256-thread workgroup).
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
// normal} (12 doubles, 16-byte aligned).
__device__ __forceinline__ bool
isect_f64(const double* __restrict__ tri, const double o[3], const double d[3], double tmin,
          double tmax, double& t_out, double& u_out, double& v_out)
{
  const double2* q = reinterpret_cast<const double2*>(tri);
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
  // The stack: capacity 128 (pushes onto a full stack are dropped), entries
  // (node index, tmin).  The first kStackF64Lds entries live in LDS, one
  // column per thread, the rest in a private array; the top entry is kept in
  // registers and reloaded only after a pop empties them (its reload is
  // issued after the popped node's fetch, so the two latencies overlap).
  __shared__ uint32_t s_node[kStackF64Lds * kTraceF64Block];
  __shared__ double s_tmin[kStackF64Lds * kTraceF64Block];
  uint32_t spill_n[kStackF64 - kStackF64Lds];
  double spill_t[kStackF64 - kStackF64Lds];
  auto stk_store = [&](int e, uint32_t node, double tm) {
    if (e < kStackF64Lds) {
      s_node[e * kTraceF64Block + threadIdx.x] = node;
      s_tmin[e * kTraceF64Block + threadIdx.x] = tm;
    } else {
      spill_n[e - kStackF64Lds] = node;
      spill_t[e - kStackF64Lds] = tm;
    }
  };
  auto stk_load = [&](int e, uint32_t& node, double& tm) {
    if (e < kStackF64Lds) {
      node = s_node[e * kTraceF64Block + threadIdx.x];
      tm = s_tmin[e * kTraceF64Block + threadIdx.x];
    } else {
      node = spill_n[e - kStackF64Lds];
      tm = spill_t[e - kStackF64Lds];
    }
  };
  uint32_t top_n = 0;
  double top_t = 0.0;
  int sp = 0;
  bool stale = false; // the registers hold an entry already popped: reload before use
  auto push = [&](uint32_t node, double tm) {
    if (sp >= kStackF64)
      return;
    if (sp > 0 && !stale)
      stk_store(sp - 1, top_n, top_t);
    stale = false;
    top_n = node;
    top_t = tm;
    sp++;
  };
  uint64_t best = ~uint64_t(0);
  double best_t = INFINITY, best_u = 0.0, best_v = 0.0;
  // The root is pushed and popped at once; a child that would be pushed and
  // popped again straight away (the near one of two, or a single hit child)
  // is taken directly, when its push would have found room, so the visit
  // order and the drop-when-full rule are the reference's.
  uint32_t ni = 0u;
  bool have = true;
  for (;;) {
    bool visit = true;
    if (!have) {
      if (sp == 0)
        break;
      if (stale)
        stk_load(sp - 1, top_n, top_t);
      ni = top_n;
      visit = !(best_t < top_t);
      sp--;
      stale = sp > 0;
    }
    have = false;
    if (!visit || uint64_t(ni) + 1 >= n)
      continue; // pruned, or (only for a corrupt node array) out of range
    const uint64_t lr[2] = { nodes[ni].left, nodes[ni].right };
    if (stale) { // reload the top while the node is in flight
      stk_load(sp - 1, top_n, top_t);
      stale = false;
    }
    bool hit2[2] = { false, false };
    double tn2[2] = { 0.0, 0.0 };
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const uint64_t ch = lr[side];
      const uint64_t idx = ch & ~kLeafBit64;
      if ((ch & kLeafBit64) ? idx >= n : idx + 1 >= n)
        continue; // out-of-range child: only possible for a corrupt node array
      if (ch & kLeafBit64) {
        double t, u, v;
        if (isect_f64(tris + 12 * idx, o, d, rtmin, rtmax, t, u, v)) {
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
        const double2* bp = reinterpret_cast<const double2*>(&nodes[idx]);
        const double2 b0 = bp[0], b1 = bp[1], b2 = bp[2];
        const double lo[3] = { b0.x, b0.y, b1.x };
        const double hi[3] = { b1.y, b2.x, b2.y };
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
    uint32_t next_n;
    double next_t;
    if (hit2[0] && hit2[1]) {
      const bool left_near = tn2[0] < tn2[1];
      push(uint32_t(left_near ? lr[1] : lr[0]), left_near ? tn2[1] : tn2[0]);
      next_n = uint32_t(left_near ? lr[0] : lr[1]);
      next_t = left_near ? tn2[0] : tn2[1];
    } else if (hit2[0] || hit2[1]) {
      next_n = uint32_t(hit2[0] ? lr[0] : lr[1]);
      next_t = hit2[0] ? tn2[0] : tn2[1];
    } else {
      continue;
    }
    // push-then-pop of next_n: its push needs room; its pop is then pruned
    // when a closer hit is already known
    if (sp < kStackF64 && !(best_t < next_t)) {
      ni = next_n;
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
