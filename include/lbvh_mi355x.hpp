// lbvh_mi355x.hpp — drop-in C++ surface of tay10r/lbvh over the MI355X engine.
//
// Replace `#include <lbvh.h>` with `#include <lbvh_mi355x.hpp>` and link
// liblbvh_mi355x.so.  The names, layouts and call shapes of the reference
// API are kept (lbvh.h:62-558 types, 1793-1942 builder/traverser,
// 1946-2080 triangle plugin); the work runs on a gfx950 GPU:
//
//   lbvh::builder<float, Sched>::operator()(prims, count, converter)
//     * triangle<float> + triangle_aabb_converter<float>: the triangles go to
//       the device as they are (lbvh_build_triangles);
//     * any other primitive/converter: the converter is evaluated on the host
//       through the caller's scheduler (as the reference calls it,
//       lbvh.h:1282-1293) and the boxes are built on the device
//       (lbvh_build_aabbs);
//     * returns a host lbvh::bvh<float> whose node array is byte-identical
//       to the reference's (REFERENCE leaf order, 30-bit codes) unless
//       lbvh::gpu::options select otherwise.
//   lbvh::traverser<float, P, Info>::operator()(ray, intersector): per-ray
//     closest hit over the host node array with the reference's visit order.
//   lbvh::gpu::device_tree: a device-resident tree for batched closest-hit
//     of triangle scenes (lbvh_trace_closest), refit and node download.
//
//   lbvh::builder<double, Sched>::operator()(prims, count, converter): the
//     converter is evaluated on the host through the caller's scheduler and
//     the FP64 boxes are built on the device (lbvh_build_aabbs_f64: u64
//     codes, FP64 arithmetic, 64-byte node<double>), byte-identical to the
//     reference's builder<double>.
//   lbvh::gpu::trace_closest(bvh<double>, triangle<double>*, ...): the
//     traverser<double> over a batch of rays on the device
//     (lbvh_trace_closest_f64), equal to the host traverser bit for bit.
#pragma once

#include "lbvh_mi355x.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace lbvh {

using size_type = std::size_t;

template<size_type type_size>
struct associated_types;

template<>
struct associated_types<4>
{
  using uint_type = std::uint32_t;
  using int_type = std::int32_t;
  using float_type = float;
};

template<>
struct associated_types<8>
{
  using uint_type = std::uint64_t;
  using int_type = std::int64_t;
  using float_type = double;
};

// ---- scheduler plugin (host work division) ---------------------------------

struct work_division
{
  size_type idx;
  size_type max;
};

class single_thread_scheduler final
{
public:
  template<typename task_type, typename... arg_types>
  void operator()(task_type task, arg_types... args) noexcept
  {
    task(work_division{ 0, 1 }, args...);
  }
  size_type max_threads() const noexcept { return 1; }
};

class naive_thread_scheduler final
{
  size_type count_;

public:
  naive_thread_scheduler(size_type max_threads_ = std::thread::hardware_concurrency())
    : count_(max_threads_ ? max_threads_ : 1)
  {}
  template<typename task_type, typename... arg_types>
  void operator()(task_type task, arg_types... args)
  {
    std::vector<std::thread> pool;
    pool.reserve(count_ - 1);
    for (size_type i = 0; i + 1 < count_; i++)
      pool.emplace_back(task, work_division{ i, count_ }, args...);
    task(work_division{ count_ - 1, count_ }, args...);
    for (auto& t : pool)
      t.join();
  }
  size_type max_threads() const noexcept { return count_; }
};

using default_scheduler = naive_thread_scheduler;

// ---- value types (layouts of lbvh.h:177-255, 384-427) --------------------

template<typename int_type>
constexpr int_type
highest_bit() noexcept
{
  return int_type(1) << (sizeof(int_type) * 8 - 1);
}

template<typename T>
struct vec2 final
{
  T x;
  T y;
};

template<typename T>
struct vec3 final
{
  T x;
  T y;
  T z;
};

template<typename T>
struct aabb final
{
  vec3<T> min;
  vec3<T> max;
};

template<typename T>
struct alignas(sizeof(T) * 8) node final
{
  using index_type = typename associated_types<sizeof(T)>::uint_type;
  aabb<T> box;
  index_type left;
  index_type right;
  constexpr index_type left_leaf_index() const noexcept
  {
    return left & (highest_bit<index_type>() - 1);
  }
  constexpr index_type right_leaf_index() const noexcept
  {
    return right & (highest_bit<index_type>() - 1);
  }
  constexpr bool left_is_leaf() const noexcept { return left & highest_bit<index_type>(); }
  constexpr bool right_is_leaf() const noexcept { return right & highest_bit<index_type>(); }
};

static_assert(sizeof(node<float>) == 32, "node<float> must match the engine's node layout");
static_assert(sizeof(node<double>) == 64, "node<double> must match the engine's FP64 node layout");

template<typename T>
class bvh final
{
public:
  using node_type = node<T>;
  using node_vec = std::vector<node_type>;
  bvh(node_vec&& nodes_)
    : nodes(std::move(nodes_))
  {}
  auto begin() const noexcept { return nodes.begin(); }
  auto end() const noexcept { return nodes.end(); }
  auto size() const noexcept { return nodes.size(); }
  const node_type& at(size_type i) const { return nodes.at(i); }
  const node_type& operator[](size_type i) const noexcept { return nodes[i]; }
  const node_type* data() const noexcept { return nodes.data(); }

private:
  node_vec nodes;
};

template<typename T>
struct ray final
{
  using vec_type = vec3<T>;
  vec_type pos;
  vec_type dir;
  T tmin = 0;
  T tmax = std::numeric_limits<T>::infinity();
};

static_assert(sizeof(ray<float>) == 32, "ray<float> must match the engine's ray layout");
static_assert(sizeof(ray<double>) == 64, "ray<double> must match the engine's FP64 ray layout");

template<typename T, typename intersection_info>
struct intersection_container final
{
  using index = typename associated_types<sizeof(T)>::uint_type;
  index primitive = std::numeric_limits<index>::max();
  intersection_info info;
  intersection_container() = default;
  intersection_container(index prim, intersection_info&& i)
    : primitive(prim)
    , info(std::move(i))
  {}
  constexpr operator bool() const noexcept
  {
    return primitive != std::numeric_limits<index>::max();
  }
};

// ---- math helpers (semantics of lbvh.h:560-905) ----------------------------

namespace math {

template<typename T>
constexpr T
min(T a, T b) noexcept
{
  return (a < b) ? a : b;
}
template<typename T>
constexpr T
max(T a, T b) noexcept
{
  return (a > b) ? a : b;
}
template<typename T>
vec3<T>
min(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z) };
}
template<typename T>
vec3<T>
max(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z) };
}
template<typename T>
constexpr vec3<T>
cross(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { (a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x) };
}
template<typename T>
constexpr T
dot(const vec3<T>& a, const vec3<T>& b) noexcept
{
  T out = 0;
  out += a.x * b.x;
  out += a.y * b.y;
  out += a.z * b.z;
  return out;
}
template<typename T>
vec3<T>
operator+(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { a.x + b.x, a.y + b.y, a.z + b.z };
}
template<typename T>
vec3<T>
operator-(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { a.x - b.x, a.y - b.y, a.z - b.z };
}
template<typename T>
vec3<T>
operator-(const vec3<T>& a) noexcept
{
  return { -a.x, -a.y, -a.z };
}
template<typename T>
vec3<T>
operator*(const vec3<T>& a, T b) noexcept
{
  return { a.x * b, a.y * b, a.z * b };
}
template<typename T>
vec2<T>
operator*(const vec2<T>& a, T b) noexcept
{
  return { a.x * b, a.y * b };
}
template<typename T>
vec2<T>
operator+(const vec2<T>& a, const vec2<T>& b) noexcept
{
  return { a.x + b.x, a.y + b.y };
}
template<typename T>
vec3<T>
hadamard_mul(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { a.x * b.x, a.y * b.y, a.z * b.z };
}
template<typename T>
vec3<T>
hadamard_div(const vec3<T>& a, const vec3<T>& b) noexcept
{
  return { a.x / b.x, a.y / b.y, a.z / b.z };
}
template<typename T>
vec3<T>
reciprocal(const vec3<T>& v) noexcept
{
  return { 1 / v.x, 1 / v.y, 1 / v.z };
}
template<typename T>
vec3<T>
normalize(const vec3<T>& v) noexcept
{
  using std::sqrt;
  T l_inv = T(1) / sqrt(dot(v, v));
  return v * l_inv;
}

} // namespace math

// ---- GPU session --------------------------------------------------------------

namespace gpu {

// Build options of the drop-in builder (defaults reproduce lbvh.h).
struct options
{
  std::uint32_t code_bits = 30;                    // 30 or 63
  std::uint32_t leaf_order = LBVH_LEAF_REFERENCE; // or LBVH_LEAF_PRIMITIVE_ID
};

class error : public std::runtime_error
{
public:
  error(lbvh_status s, const std::string& what)
    : std::runtime_error(std::string(lbvh_status_string(s)) + ": " + what)
    , status(s)
  {}
  lbvh_status status;
};

// The process-wide engine context used by the drop-in builder.
class session
{
public:
  static session& get()
  {
    static session s;
    return s;
  }
  lbvh_ctx* ctx()
  {
    std::call_once(once_, [this] {
      lbvh_ctx* c = nullptr;
      lbvh_status st = lbvh_ctx_create(device_, &c);
      if (st != LBVH_OK)
        throw error(st, "lbvh_ctx_create");
      ctx_.reset(c);
    });
    return ctx_.get();
  }
  void check(lbvh_status st, const char* what)
  {
    if (st != LBVH_OK)
      throw error(st, std::string(what) + ": " + lbvh_ctx_last_error(ctx_.get()));
  }
  options opts;
  int device_ = 0;

private:
  struct ctx_del
  {
    void operator()(lbvh_ctx* c) const { lbvh_ctx_destroy(c); }
  };
  std::once_flag once_;
  std::unique_ptr<lbvh_ctx, ctx_del> ctx_;
};

inline options&
default_options()
{
  return session::get().opts;
}

} // namespace gpu

template<typename T>
class triangle;
template<typename T>
class triangle_aabb_converter;

// ---- builder -------------------------------------------------------------------

template<typename scalar_type, typename task_scheduler = default_scheduler>
class builder final
{
  static_assert(std::is_same<scalar_type, float>::value ||
                  std::is_same<scalar_type, double>::value,
                "lbvh::builder serves float and double");
  task_scheduler scheduler;

  // boxes[i] = converter(primitives[i]), split over the caller's scheduler
  // as the reference divides its loops (lbvh.h:1282-1293).
  template<typename primitive, typename aabb_converter>
  std::vector<aabb<scalar_type>> convert(const primitive* primitives, size_type count,
                                         const aabb_converter& converter)
  {
    std::vector<aabb<scalar_type>> boxes(count);
    auto task = [&](const work_division& div) {
      size_type chunk = count / div.max;
      size_type b = chunk * div.idx;
      size_type e = (div.idx + 1 == div.max) ? count : b + chunk;
      for (size_type i = b; i < e; i++)
        boxes[i] = converter(primitives[i]);
    };
    scheduler(task);
    return boxes;
  }

public:
  using bvh_type = bvh<scalar_type>;
  using node_type = node<scalar_type>;
  using node_vec = std::vector<node_type>;

  builder(task_scheduler s = task_scheduler())
    : scheduler(s)
  {}

  template<typename primitive, typename aabb_converter>
  bvh_type operator()(const primitive* primitives, size_type count,
                      const aabb_converter& converter)
  {
    auto& S = gpu::session::get();
    lbvh_ctx* c = S.ctx();
    lbvh_build_opts o;
    lbvh_build_opts_init(&o);
    o.code_bits = S.opts.code_bits;
    o.leaf_order = S.opts.leaf_order;
    if (count < 2)
      throw std::invalid_argument("lbvh::builder: a tree needs at least 2 primitives");
    if constexpr (std::is_same<scalar_type, double>::value) {
      // code width follows the scalar type: u64 codes, as the reference
      auto boxes = convert(primitives, count, converter);
      node_vec nodes(count - 1);
      S.check(lbvh_build_aabbs_f64(c, reinterpret_cast<const double*>(boxes.data()), count, &o,
                                   nodes.data()),
              "lbvh_build_aabbs_f64");
      return bvh_type(std::move(nodes));
    } else {
      lbvh_bvh* h = nullptr;
      if (std::is_same<primitive, triangle<float>>::value &&
          std::is_same<aabb_converter, triangle_aabb_converter<float>>::value) {
        S.check(lbvh_build_triangles(c, reinterpret_cast<const float*>(primitives), count, &o, &h),
                "lbvh_build_triangles");
      } else {
        auto boxes = convert(primitives, count, converter);
        S.check(lbvh_build_aabbs(c, reinterpret_cast<const float*>(boxes.data()), count, &o, &h),
                "lbvh_build_aabbs");
      }
      node_vec nodes(count - 1);
      lbvh_status st =
        lbvh_bvh_download_nodes(c, h, nodes.data(), nodes.size() * sizeof(node_type));
      lbvh_bvh_destroy(h);
      S.check(st, "lbvh_bvh_download_nodes");
      return bvh_type(std::move(nodes));
    }
  }
};

// ---- traverser (host, per ray) ---------------------------------------------------

template<typename scalar_type, typename primitive_type, typename intersection_info>
class traverser final
{
  const bvh<scalar_type>& bvh_;
  const primitive_type* primitives;

public:
  using intersection = intersection_container<scalar_type, intersection_info>;
  using ray_type = ray<scalar_type>;

  constexpr traverser(const bvh<scalar_type>& b, const primitive_type* p) noexcept
    : bvh_(b)
    , primitives(p)
  {}

  // Visit order of the reference traverser: root first with tmin = +inf;
  // skip entries farther than the closest hit; leaves intersected at once,
  // left then right; internal children pushed far first; 128-entry stack
  // that drops pushes when full.
  template<typename intersector_type>
  intersection operator()(const ray_type& r, const intersector_type& isect) const noexcept
  {
    using idx_t = typename node<scalar_type>::index_type;
    const scalar_type inf = std::numeric_limits<scalar_type>::infinity();
    struct entry
    {
      idx_t node;
      scalar_type tmin;
    } stack[128];
    size_type sp = 0;
    stack[sp++] = entry{ 0, inf };
    const vec3<scalar_type> rcp = math::reciprocal(r.dir);
    const vec3<scalar_type> ipos = math::hadamard_mul(math::operator-(r.pos), rcp);
    const bool px = r.dir.x > 0, py = r.dir.y > 0, pz = r.dir.z > 0;
    auto box_test = [&](const aabb<scalar_type>& b, scalar_type& tmin_out) {
      scalar_type tn0 = ((px ? b.min.x : b.max.x) * rcp.x) + ipos.x;
      scalar_type tn1 = ((py ? b.min.y : b.max.y) * rcp.y) + ipos.y;
      scalar_type tn2 = ((pz ? b.min.z : b.max.z) * rcp.z) + ipos.z;
      scalar_type tf0 = ((px ? b.max.x : b.min.x) * rcp.x) + ipos.x;
      scalar_type tf1 = ((py ? b.max.y : b.min.y) * rcp.y) + ipos.y;
      scalar_type tf2 = ((pz ? b.max.z : b.min.z) * rcp.z) + ipos.z;
      scalar_type tmin = math::max(math::max(tn0, tn1), tn2);
      scalar_type tmax = math::min(math::min(tf0, tf1), tf2);
      tmin_out = tmin;
      return tmax >= math::max(scalar_type(0), tmin);
    };
    auto push = [&](idx_t n, scalar_type t) {
      if (sp < 128)
        stack[sp++] = entry{ n, t };
    };
    intersection closest;
    while (sp) {
      entry e = stack[--sp];
      if (closest.info < e.tmin)
        continue;
      const auto& nd = bvh_[e.node];
      bool lh = false, rh = false;
      scalar_type lt = 0, rt = 0;
      if (nd.left_is_leaf()) {
        auto i = nd.left_leaf_index();
        auto info = isect(primitives[i], r);
        if (info < closest.info)
          closest = intersection(i, std::move(info));
      } else {
        lh = box_test(bvh_[nd.left].box, lt);
      }
      if (nd.right_is_leaf()) {
        auto i = nd.right_leaf_index();
        auto info = isect(primitives[i], r);
        if (info < closest.info)
          closest = intersection(i, std::move(info));
      } else {
        rh = box_test(bvh_[nd.right].box, rt);
      }
      if (lh && rh) {
        if (lt < rt) {
          push(nd.right, rt);
          push(nd.left, lt);
        } else {
          push(nd.left, lt);
          push(nd.right, rt);
        }
      } else if (lh) {
        push(nd.left, lt);
      } else if (rh) {
        push(nd.right, rt);
      }
    }
    return closest;
  }
};

// ---- triangle plugin (lbvh.h:1946-2080 interface) ---------------------------------

template<typename T>
class triangle_intersector;

template<typename T>
class triangle final
{
public:
  using vec3_type = vec3<T>;
  constexpr triangle() = default;
  constexpr triangle(const vec3_type& a, const vec3_type& b, const vec3_type& c)
    : p0(a)
    , e0(math::operator-(a, b))
    , e1(math::operator-(c, a))
    , normal(math::cross(e0, e1))
  {}

private:
  friend triangle_intersector<T>;
  friend triangle_aabb_converter<T>;
  vec3_type p0;
  vec3_type e0;
  vec3_type e1;
  vec3_type normal;
};

static_assert(sizeof(triangle<float>) == 48, "triangle<float> must match the engine's layout");
static_assert(sizeof(triangle<double>) == 96, "triangle<double> must match the engine's layout");

template<typename T>
class triangle_aabb_converter final
{
public:
  using box_type = aabb<T>;
  using triangle_type = triangle<T>;
  box_type operator()(const triangle_type& t) const noexcept
  {
    using namespace math;
    const auto p1 = t.p0 - t.e0;
    const auto p2 = t.p0 + t.e1;
    return box_type{ min(min(t.p0, p1), p2), max(max(t.p0, p1), p2) };
  }
};

template<typename T>
struct triangle_intersection final
{
  T distance = std::numeric_limits<T>::infinity();
  vec2<T> uv;
  constexpr bool operator<(const triangle_intersection& o) const noexcept
  {
    return distance < o.distance;
  }
  constexpr bool operator<(T d) const noexcept { return distance < d; }
};

template<typename T>
class triangle_intersector final
{
public:
  using intersection_type = triangle_intersection<T>;
  using ray_type = ray<T>;
  intersection_type operator()(const triangle<T>& tri, const ray_type& r) const noexcept
  {
    using namespace math;
    const vec3<T> c = tri.p0 - r.pos;
    const vec3<T> rr = math::cross(r.dir, c);
    const T inv_det = 1 / math::dot(tri.normal, r.dir);
    const T u = dot(rr, tri.e1) * inv_det;
    const T v = dot(rr, tri.e0) * inv_det;
    const T w = T(1) - u - v;
    if ((u >= 0) && (v >= 0) && (w >= 0)) {
      const T t = dot(tri.normal, c) * inv_det;
      if ((t >= r.tmin) && (t <= r.tmax))
        return intersection_type{ t, { u, v } };
    }
    return intersection_type{};
  }
};

// ---- device-resident trees (batched closest hit) ---------------------------------

namespace gpu {

// A triangle tree kept on the GPU: build, refit, batched closest hit, and
// download as the reference's bvh<float>.
class device_tree
{
public:
  device_tree(const triangle<float>* tris, size_type count, const options& o = options())
  {
    auto& S = session::get();
    lbvh_build_opts bo;
    lbvh_build_opts_init(&bo);
    bo.code_bits = o.code_bits;
    bo.leaf_order = o.leaf_order;
    lbvh_bvh* h = nullptr;
    S.check(lbvh_build_triangles(S.ctx(), reinterpret_cast<const float*>(tris), count, &bo, &h),
            "lbvh_build_triangles");
    h_.reset(h);
    n_ = count;
  }
  void refit(const triangle<float>* tris)
  {
    auto& S = session::get();
    S.check(lbvh_refit_triangles(S.ctx(), h_.get(), reinterpret_cast<const float*>(tris)),
            "lbvh_refit_triangles");
  }
  // Closest hit of every ray; prim = UINT32_MAX on a miss.
  void trace(const ray<float>* rays, size_type nrays, std::uint32_t* prim, float* t,
             vec2<float>* uv = nullptr, int mode = LBVH_ISECT_REFERENCE) const
  {
    auto& S = session::get();
    S.check(lbvh_trace_closest(S.ctx(), h_.get(), reinterpret_cast<const float*>(rays), nrays,
                               mode, prim, t, reinterpret_cast<float*>(uv)),
            "lbvh_trace_closest");
  }
  bvh<float> download() const
  {
    auto& S = session::get();
    std::vector<node<float>> nodes(n_ - 1);
    S.check(lbvh_bvh_download_nodes(S.ctx(), h_.get(), nodes.data(), nodes.size() * 32),
            "lbvh_bvh_download_nodes");
    return bvh<float>(std::move(nodes));
  }
  lbvh_bvh* handle() const { return h_.get(); }

private:
  struct del
  {
    void operator()(lbvh_bvh* b) const { lbvh_bvh_destroy(b); }
  };
  std::unique_ptr<lbvh_bvh, del> h_;
  size_type n_ = 0;
};

// traverser<double, triangle<double>, triangle_intersection<double>> applied
// to every ray of a batch on the device (lbvh_trace_closest_f64): the closest
// hit over a builder<double> tree of `count` triangles, equal bit for bit to
// the host traverser's.  prim = UINT64_MAX on a miss; uv may be null; mode
// LBVH_ISECT_REFERENCE (first hit wins ties) or LBVH_ISECT_REFERENCE_LOWID.
inline void
trace_closest(const bvh<double>& b, const triangle<double>* tris, size_type count,
              const ray<double>* rays, size_type nrays, std::uint64_t* prim, double* t,
              vec2<double>* uv = nullptr, int mode = LBVH_ISECT_REFERENCE)
{
  auto& S = session::get();
  if (b.size() + 1 != count)
    throw std::invalid_argument("lbvh::gpu::trace_closest: the tree has count - 1 nodes");
  S.check(lbvh_trace_closest_f64(S.ctx(), b.data(), count,
                                 reinterpret_cast<const double*>(tris),
                                 reinterpret_cast<const double*>(rays), nrays, mode, prim, t,
                                 reinterpret_cast<double*>(uv)),
          "lbvh_trace_closest_f64");
}

} // namespace gpu

} // namespace lbvh
