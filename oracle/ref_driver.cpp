// ORACLE — test infrastructure only.
//
// A thin C-callable driver around the real tay10r/lbvh header.  It is
// compiled from the reference's own lbvh.h (include path given by the recipe
// in oracle/build_oracle.py) into oracle/_ref/libreflbvh.so, which is never
// committed.  It pins the clean-room restatement in lbvh_oracle.c and serves
// as the "reference" CPU baseline.  Nothing in the product links it.
//
// What it runs:
//   * 30-bit codes, REFERENCE leaves: lbvh::builder<float> as shipped
//     (lbvh.h:1796-1822).
//   * 63-bit codes over FP32 triangles: the reference cannot express this
//     through builder<float> (its code width follows sizeof(scalar)), so the
//     pipeline is composed from the reference's own pieces:
//     detail::centroid_bounds_kernel, detail::center_of/size_of,
//     detail::morton_domain<8>, detail::morton_encoder<8>,
//     detail::space_filling_curve<uint64_t>::sort and
//     detail::builder_kernel<uint64_t, float>; only the 20-line encode loop
//     and the fit (protected in a final class, lbvh.h:347-352) are restated.
//   * PRIMITIVE_ID leaves: the primitives are first permuted into stable
//     (code, id) order and built with the same reference pipeline; a leaf at
//     sorted position k then names primitive perm[k].
//   * builder<double> (lbvh.h:549-558): as shipped, over triangle<double>
//     with triangle_aabb_converter<double> and over plain aabb<double> with
//     an identity converter; PRIMITIVE_ID leaves as above.
//   * Traversal: lbvh::traverser with triangle_intersector (lbvh.h:1874-1942,
//     2055-2075); the lowest-id tie-break wraps the primitive and its
//     intersection record, leaving the traverser unmodified.
#include <lbvh.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

using tri_t = lbvh::triangle<float>;
using cvt_t = lbvh::triangle_aabb_converter<float>;
using node_t = lbvh::node<float>;
using sched_t = lbvh::naive_thread_scheduler;

static_assert(sizeof(tri_t) == 48, "triangle<float> layout");
static_assert(sizeof(node_t) == 32, "node<float> layout");

std::vector<tri_t>
load_tris(const float* p, uint64_t n)
{
  std::vector<tri_t> v(n);
  std::memcpy(static_cast<void*>(v.data()), p, n * sizeof(tri_t));
  return v;
}

lbvh::aabb<float>
centroid_bounds(sched_t& sched, const tri_t* prims, uint64_t n,
                const cvt_t& cvt)
{
  using kern_t = lbvh::detail::centroid_bounds_kernel<float, tri_t, cvt_t>;
  std::vector<lbvh::aabb<float>> boxes(sched.max_threads());
  kern_t k(prims, n, cvt, boxes.data());
  sched(k);
  auto b = lbvh::detail::get_empty_aabb<float>();
  for (const auto& tb : boxes)
    b = lbvh::detail::union_of(b, tb);
  return b;
}

// 64-bit-code encode over FP32 centres: morton_curve_kernel's arithmetic
// (lbvh.h:1432-1473) with morton_domain<8> and morton_encoder<8>.
std::vector<lbvh::detail::space_filling_curve<uint64_t>::entry>
encode63(const tri_t* prims, uint64_t n, const lbvh::aabb<float>& cb,
         const cvt_t& cvt, unsigned threads)
{
  using entry = lbvh::detail::space_filling_curve<uint64_t>::entry;
  std::vector<entry> e(n);
  const float dom = float(lbvh::detail::morton_domain<8>::value());
  auto ext = lbvh::detail::size_of(cb);
  const lbvh::vec3<float> scale{ dom / (ext.x + 1.0f), dom / (ext.y + 1.0f),
                                 dom / (ext.z + 1.0f) };
  auto work = [&](unsigned t) {
    lbvh::detail::morton_encoder<8> enc;
    uint64_t chunk = n / threads;
    uint64_t b = chunk * t, end = (t + 1 == threads) ? n : b + chunk;
    for (uint64_t i = b; i < end; i++) {
      auto c = lbvh::detail::center_of(cvt(prims[i]));
      float qx = (c.x - cb.min.x) * scale.x;
      float qy = (c.y - cb.min.y) * scale.y;
      float qz = (c.z - cb.min.z) * scale.z;
      e[i] = entry{ enc(uint64_t(qx), uint64_t(qy), uint64_t(qz)), i };
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < threads; t++)
    th.emplace_back(work, t);
  work(0);
  for (auto& x : th)
    x.join();
  return e;
}

// builder::fit_boxes (lbvh.h:1827-1866), restated because it is protected.
void
fit(std::vector<node_t>& nodes, const tri_t* prims, const cvt_t& cvt)
{
  std::vector<size_t> idx;
  idx.reserve(nodes.size());
  idx.push_back(0);
  for (size_t i = 0; i < idx.size(); i++) {
    auto j = idx[i];
    if (!nodes[j].left_is_leaf())
      idx.push_back(nodes[j].left);
    if (!nodes[j].right_is_leaf())
      idx.push_back(nodes[j].right);
  }
  for (size_t i = idx.size(); i > 0; i--) {
    auto& nd = nodes[idx[i - 1]];
    if (nd.left_is_leaf())
      nd.box = cvt(prims[nd.left_leaf_index()]);
    else
      nd.box = nodes[nd.left].box;
    if (nd.right_is_leaf())
      nd.box = lbvh::detail::union_of(nd.box, cvt(prims[nd.right_leaf_index()]));
    else
      nd.box = lbvh::detail::union_of(nd.box, nodes[nd.right].box);
  }
}

struct phase_ms
{
  double encode = 0, sort = 0, tree = 0, fit = 0;
};

using clk = std::chrono::steady_clock;
double
ms_since(clk::time_point t0)
{
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

// The 63-bit pipeline on `prims` (REFERENCE leaves).
std::vector<node_t>
build63(const tri_t* prims, uint64_t n, unsigned threads, phase_ms* ph,
        std::vector<uint64_t>* sorted_codes, lbvh::aabb<float>* bounds_out)
{
  sched_t sched(threads);
  cvt_t cvt;
  auto t0 = clk::now();
  auto cb = centroid_bounds(sched, prims, n, cvt);
  auto entries = encode63(prims, n, cb, cvt, threads);
  if (ph)
    ph->encode = ms_since(t0);
  t0 = clk::now();
  lbvh::detail::space_filling_curve<uint64_t> curve(std::move(entries));
  curve.sort();
  if (ph)
    ph->sort = ms_since(t0);
  t0 = clk::now();
  std::vector<node_t> nodes(curve.size() - 1);
  lbvh::detail::builder_kernel<uint64_t, float> bk(curve, nodes.data());
  sched(bk);
  if (ph)
    ph->tree = ms_since(t0);
  t0 = clk::now();
  fit(nodes, prims, cvt);
  if (ph)
    ph->fit = ms_since(t0);
  if (sorted_codes) {
    sorted_codes->resize(n);
    for (uint64_t i = 0; i < n; i++)
      (*sorted_codes)[i] = curve[i].code;
  }
  if (bounds_out)
    *bounds_out = cb;
  return nodes;
}

// The 30-bit pipeline through the reference's own pieces with phase timing.
std::vector<node_t>
build30_phased(const tri_t* prims, uint64_t n, unsigned threads, phase_ms* ph,
               std::vector<uint32_t>* sorted_codes)
{
  sched_t sched(threads);
  cvt_t cvt;
  auto t0 = clk::now();
  lbvh::detail::morton_curve_builder<float, sched_t> cb(sched);
  auto curve = cb(prims, n, cvt);
  ph->encode = ms_since(t0);
  t0 = clk::now();
  curve.sort();
  ph->sort = ms_since(t0);
  t0 = clk::now();
  std::vector<node_t> nodes(curve.size() - 1);
  lbvh::detail::builder_kernel<uint32_t, float> bk(curve, nodes.data());
  sched(bk);
  ph->tree = ms_since(t0);
  t0 = clk::now();
  fit(nodes, prims, cvt);
  ph->fit = ms_since(t0);
  if (sorted_codes) {
    sorted_codes->resize(n);
    for (uint64_t i = 0; i < n; i++)
      (*sorted_codes)[i] = curve[i].code;
  }
  return nodes;
}

// Unsorted codes of the primitives, as the reference computes them.
template<typename code_t>
std::vector<std::pair<code_t, uint64_t>>
curve_pairs(const tri_t* prims, uint64_t n, int code_bits, unsigned threads)
{
  std::vector<std::pair<code_t, uint64_t>> out(n);
  sched_t sched(threads);
  cvt_t cvt;
  if (code_bits == 30) {
    lbvh::detail::morton_curve_builder<float, sched_t> cb(sched);
    auto curve = cb(prims, n, cvt);
    for (uint64_t i = 0; i < n; i++)
      out[i] = { code_t(curve[i].code), uint64_t(curve[i].primitive) };
  } else {
    auto b = centroid_bounds(sched, prims, n, cvt);
    auto e = encode63(prims, n, b, cvt, threads);
    for (uint64_t i = 0; i < n; i++)
      out[i] = { code_t(e[i].code), uint64_t(e[i].primitive) };
  }
  return out;
}

// Stable (code, id) permutation.
std::vector<uint32_t>
stable_perm(const tri_t* prims, uint64_t n, int code_bits, unsigned threads)
{
  auto pairs = curve_pairs<uint64_t>(prims, n, code_bits, threads);
  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; i++)
    perm[i] = uint32_t(i);
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
    return pairs[a].first < pairs[b].first;
  });
  return perm;
}

struct lowid_tri
{
  tri_t t;
  uint32_t id;
};

struct lowid_hit
{
  lbvh::triangle_intersection<float> h;
  uint32_t id = UINT32_MAX;
  bool operator<(const lowid_hit& o) const noexcept
  {
    return (h.distance < o.h.distance) ||
           (h.distance == o.h.distance && id < o.id);
  }
  bool operator<(float d) const noexcept { return h.distance < d; }
};

// traverser<double> with equal-distance ties to the lowest id: the id rides
// with the primitive, and the intersection orders by (distance, id) while
// pruning compares the distance alone.
struct lowid_tri64
{
  lbvh::triangle<double> t;
  uint64_t id;
};

struct lowid_hit64
{
  lbvh::triangle_intersection<double> h;
  uint64_t id = UINT64_MAX;
  bool operator<(const lowid_hit64& o) const noexcept
  {
    return (h.distance < o.h.distance) || (h.distance == o.h.distance && id < o.id);
  }
  bool operator<(double d) const noexcept { return h.distance < d; }
};

// builder<double>: the converter of plain FP64 boxes (the generic
// primitive/converter path with P = aabb<double>).
struct box_cvt_f64
{
  lbvh::aabb<double> operator()(const lbvh::aabb<double>& b) const noexcept { return b; }
};

using node64_t = lbvh::node<double>;
static_assert(sizeof(node64_t) == 64, "node<double> layout");
static_assert(sizeof(lbvh::triangle<double>) == 96, "triangle<double> layout");

// lbvh::builder<double> as shipped (lbvh.h:1796-1822); PRIMITIVE_ID leaves
// as for FP32: build over the primitives in stable (code, id) order, then
// name the primitive behind each leaf.
template<typename P, typename C>
std::vector<node64_t>
build_f64(const P* prims, uint64_t n, int leaf_order, unsigned th, const C& cvt)
{
  std::vector<P> permuted;
  std::vector<uint32_t> perm;
  const P* src = prims;
  if (leaf_order == 1) {
    sched_t sched(th);
    lbvh::detail::morton_curve_builder<double, sched_t> cb(sched);
    auto curve = cb(prims, n, cvt);
    perm.resize(n);
    for (uint64_t i = 0; i < n; i++)
      perm[i] = uint32_t(i);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return curve[a].code < curve[b].code; });
    permuted.resize(n);
    for (uint64_t i = 0; i < n; i++)
      permuted[i] = prims[perm[i]];
    src = permuted.data();
  }
  lbvh::builder<double, sched_t> b{ sched_t(th) };
  auto bvh = b(src, n, cvt);
  std::vector<node64_t> nodes(bvh.begin(), bvh.end());
  if (leaf_order == 1) {
    const uint64_t bit = uint64_t(1) << 63;
    for (auto& nd : nodes) {
      if (nd.left_is_leaf())
        nd.left = perm[nd.left_leaf_index()] | bit;
      if (nd.right_is_leaf())
        nd.right = perm[nd.right_leaf_index()] | bit;
    }
  }
  return nodes;
}

} // namespace

extern "C" {

// Builds with the reference pipeline.  leaf_order 0 = reference (leaf =
// sorted position), 1 = primitive id.  Returns 0 on success.
int
ref_build(const float* tri48, uint64_t n, int code_bits, int leaf_order,
          int threads, void* nodes_out, void* sorted_codes_out,
          float* bounds_out, double* phase_ms_out)
{
  if (n < 2 || (code_bits != 30 && code_bits != 63))
    return -1;
  unsigned th = threads > 0 ? unsigned(threads) : 1u;
  auto prims = load_tris(tri48, n);
  std::vector<uint32_t> perm;
  if (leaf_order == 1) {
    perm = stable_perm(prims.data(), n, code_bits, th);
    std::vector<tri_t> p2(n);
    for (uint64_t i = 0; i < n; i++)
      p2[i] = prims[perm[i]];
    prims.swap(p2);
  }
  std::vector<node_t> nodes;
  phase_ms ph;
  if (code_bits == 30) {
    std::vector<uint32_t> sc;
    if (phase_ms_out) {
      nodes = build30_phased(prims.data(), n, th, &ph,
                             sorted_codes_out ? &sc : nullptr);
    } else {
      lbvh::builder<float, sched_t> b{ sched_t(th) };
      auto bvh = b(prims.data(), n, cvt_t());
      nodes.assign(bvh.begin(), bvh.end());
      if (sorted_codes_out) {
        sc.resize(n);
        auto pairs = curve_pairs<uint32_t>(prims.data(), n, 30, th);
        std::vector<uint32_t> codes(n);
        for (uint64_t i = 0; i < n; i++)
          codes[i] = pairs[i].first;
        std::sort(codes.begin(), codes.end());
        sc = codes;
      }
    }
    if (sorted_codes_out)
      std::memcpy(sorted_codes_out, sc.data(), n * 4);
  } else {
    std::vector<uint64_t> sc;
    nodes = build63(prims.data(), n, th, phase_ms_out ? &ph : nullptr,
                    sorted_codes_out ? &sc : nullptr, nullptr);
    if (sorted_codes_out)
      std::memcpy(sorted_codes_out, sc.data(), n * 8);
  }
  if (leaf_order == 1) {
    for (auto& nd : nodes) {
      if (nd.left_is_leaf())
        nd.left = perm[nd.left_leaf_index()] | 0x80000000u;
      if (nd.right_is_leaf())
        nd.right = perm[nd.right_leaf_index()] | 0x80000000u;
    }
  }
  std::memcpy(nodes_out, static_cast<const void*>(nodes.data()),
              nodes.size() * sizeof(node_t));
  if (bounds_out) {
    sched_t sched(th);
    auto b = centroid_bounds(sched, prims.data(), n, cvt_t());
    float v[6] = { b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z };
    std::memcpy(bounds_out, v, sizeof(v));
  }
  if (phase_ms_out) {
    phase_ms_out[0] = ph.encode;
    phase_ms_out[1] = ph.sort;
    phase_ms_out[2] = ph.tree;
    phase_ms_out[3] = ph.fit;
  }
  return 0;
}

// Unsorted codes in primitive order, as the reference encodes them.
int
ref_codes(const float* tri48, uint64_t n, int code_bits, int threads,
          void* codes_out)
{
  if (n < 1 || (code_bits != 30 && code_bits != 63))
    return -1;
  unsigned th = threads > 0 ? unsigned(threads) : 1u;
  auto prims = load_tris(tri48, n);
  if (code_bits == 30) {
    auto pairs = curve_pairs<uint32_t>(prims.data(), n, 30, th);
    auto* o = static_cast<uint32_t*>(codes_out);
    for (uint64_t i = 0; i < n; i++)
      o[pairs[i].second] = pairs[i].first;
  } else {
    auto pairs = curve_pairs<uint64_t>(prims.data(), n, 63, th);
    auto* o = static_cast<uint64_t*>(codes_out);
    for (uint64_t i = 0; i < n; i++)
      o[pairs[i].second] = pairs[i].first;
  }
  return 0;
}

// Closest hit with the reference traverser over a node array whose leaves
// index `tri48`.  mode 0: reference tie rule; 1: lowest id on equal t.
int
ref_trace(const void* nodes_in, uint64_t n, const float* tri48,
          const float* rays, uint64_t nrays, int mode, int threads,
          uint32_t* prim, float* t, float* uv)
{
  if (n < 2)
    return -1;
  std::vector<node_t> nv(n - 1);
  std::memcpy(static_cast<void*>(nv.data()), nodes_in, (n - 1) * 32);
  lbvh::bvh<float> bvh(std::move(nv));
  auto prims = load_tris(tri48, n);
  unsigned th = threads > 0 ? unsigned(threads) : 1u;
  std::vector<lowid_tri> lp;
  if (mode == 1) {
    lp.resize(n);
    for (uint64_t i = 0; i < n; i++)
      lp[i] = lowid_tri{ prims[i], uint32_t(i) };
  }
  auto work = [&](unsigned tid) {
    lbvh::triangle_intersector<float> isect;
    if (mode == 0) {
      lbvh::traverser<float, tri_t, lbvh::triangle_intersection<float>> tr(
        bvh, prims.data());
      for (uint64_t i = tid; i < nrays; i += th) {
        const float* r = rays + 8 * i;
        lbvh::ray<float> ray{ { r[0], r[1], r[2] }, { r[3], r[4], r[5] }, r[6],
                              r[7] };
        auto hit = tr(ray, isect);
        prim[i] = hit.primitive;
        t[i] = hit.info.distance;
        if (uv) {
          uv[2 * i] = hit ? hit.info.uv.x : 0.0f;
          uv[2 * i + 1] = hit ? hit.info.uv.y : 0.0f;
        }
      }
    } else {
      auto wrap = [&isect](const lowid_tri& p,
                           const lbvh::ray<float>& r) -> lowid_hit {
        lowid_hit h;
        h.h = isect(p.t, r);
        h.id = (h.h.distance != std::numeric_limits<float>::infinity())
                 ? p.id
                 : UINT32_MAX;
        return h;
      };
      lbvh::traverser<float, lowid_tri, lowid_hit> tr(bvh, lp.data());
      for (uint64_t i = tid; i < nrays; i += th) {
        const float* r = rays + 8 * i;
        lbvh::ray<float> ray{ { r[0], r[1], r[2] }, { r[3], r[4], r[5] }, r[6],
                              r[7] };
        auto hit = tr(ray, wrap);
        prim[i] = hit.primitive;
        t[i] = hit.info.h.distance;
        if (uv) {
          uv[2 * i] = hit ? hit.info.h.uv.x : 0.0f;
          uv[2 * i + 1] = hit ? hit.info.h.uv.y : 0.0f;
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned i = 1; i < th; i++)
    pool.emplace_back(work, i);
  work(0);
  for (auto& x : pool)
    x.join();
  return 0;
}

// builder<double> over triangle<double> records (12 doubles each: p0, e0,
// e1, normal) with triangle_aabb_converter<double>.  Returns 0 on success.
int
ref_build_f64_triangles(const double* tri96, uint64_t n, int leaf_order, int threads,
                        void* nodes_out)
{
  if (n < 2)
    return -1;
  std::vector<lbvh::triangle<double>> prims(n);
  std::memcpy(static_cast<void*>(prims.data()), tri96, n * 96);
  auto nodes = build_f64(prims.data(), n, leaf_order, threads > 0 ? unsigned(threads) : 1u,
                         lbvh::triangle_aabb_converter<double>());
  std::memcpy(nodes_out, static_cast<const void*>(nodes.data()), nodes.size() * 64);
  return 0;
}

// builder<double> over plain boxes (6 doubles each).
int
ref_build_f64_aabbs(const double* boxes48, uint64_t n, int leaf_order, int threads,
                    void* nodes_out)
{
  if (n < 2)
    return -1;
  std::vector<lbvh::aabb<double>> prims(n);
  std::memcpy(static_cast<void*>(prims.data()), boxes48, n * 48);
  auto nodes = build_f64(prims.data(), n, leaf_order, threads > 0 ? unsigned(threads) : 1u,
                         box_cvt_f64());
  std::memcpy(nodes_out, static_cast<const void*>(nodes.data()), nodes.size() * 64);
  return 0;
}

// traverser<double> + triangle_intersector<double> over a node<double> array
// (64 B each), triangle<double> records (96 B) and ray<double> records
// (8 doubles).  mode 0: as shipped (first hit wins ties); 1: lowest id on
// equal distance.  uv is 0 on a miss.  Returns 0 on success.
int
ref_trace_f64(const void* nodes_in, uint64_t n, const double* tri96, const double* rays,
              uint64_t nrays, int mode, int threads, uint64_t* prim, double* t, double* uv)
{
  if (n < 2)
    return -1;
  std::vector<node64_t> nv(n - 1);
  std::memcpy(static_cast<void*>(nv.data()), nodes_in, (n - 1) * 64);
  lbvh::bvh<double> bvh(std::move(nv));
  std::vector<lbvh::triangle<double>> prims(n);
  std::memcpy(static_cast<void*>(prims.data()), tri96, n * 96);
  std::vector<lowid_tri64> lp;
  if (mode == 1) {
    lp.resize(n);
    for (uint64_t i = 0; i < n; i++)
      lp[i] = lowid_tri64{ prims[i], i };
  }
  const unsigned th = threads > 0 ? unsigned(threads) : 1u;
  auto work = [&](unsigned tid) {
    lbvh::triangle_intersector<double> isect;
    for (uint64_t i = tid; i < nrays; i += th) {
      const double* r = rays + 8 * i;
      lbvh::ray<double> ray{ { r[0], r[1], r[2] }, { r[3], r[4], r[5] }, r[6], r[7] };
      uint64_t p;
      lbvh::triangle_intersection<double> h;
      if (mode == 0) {
        lbvh::traverser<double, lbvh::triangle<double>, lbvh::triangle_intersection<double>> tr(
          bvh, prims.data());
        auto hit = tr(ray, isect);
        p = hit.primitive;
        h = hit.info;
      } else {
        auto wrap = [&isect](const lowid_tri64& q, const lbvh::ray<double>& rr) -> lowid_hit64 {
          lowid_hit64 x;
          x.h = isect(q.t, rr);
          x.id = (x.h.distance != std::numeric_limits<double>::infinity()) ? q.id : UINT64_MAX;
          return x;
        };
        lbvh::traverser<double, lowid_tri64, lowid_hit64> tr(bvh, lp.data());
        auto hit = tr(ray, wrap);
        p = hit.primitive;
        h = hit.info.h;
      }
      prim[i] = p;
      t[i] = h.distance;
      if (uv) {
        const bool hit = p != UINT64_MAX;
        uv[2 * i] = hit ? h.uv.x : 0.0;
        uv[2 * i + 1] = hit ? h.uv.y : 0.0;
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned i = 1; i < th; i++)
    pool.emplace_back(work, i);
  work(0);
  for (auto& x : pool)
    x.join();
  return 0;
}

// Hardware threads the reference's default scheduler would use.
int
ref_hardware_concurrency(void)
{
  return int(std::thread::hardware_concurrency());
}

} // extern "C"
