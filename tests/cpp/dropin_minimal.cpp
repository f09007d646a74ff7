// Compile-and-run conformance check of the drop-in header: the call shapes
// of the reference's own users (a lambda converter over a user primitive,
// the triangle plugin, the traverser with a user intersection type) compile
// unchanged against lbvh_mi355x.hpp and run on the GPU engine.
//
// Usage: dropin_minimal <out.bin>
// Writes: u64 n, n triangles (48 B each), (n-1) nodes (32 B each),
//         u64 m, m rays (32 B), m u32 prim, m f32 t (host traverser),
//         then the 2 nodes of the 3-sphere tree, then u64 n, n FP64
//         triangles (96 B each) and their builder<double> nodes (64 B each)
//         with reference leaves, then with primitive-id leaves.
// Exits non-zero if the device-tree closest hits differ from the host ones
// (FP32 device_tree and the FP64 batch traversal alike).
#include <lbvh_mi355x.hpp>

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

struct sphere
{
  float radius;
  float pos[3];
};

struct sphere_hit
{
  float tmin = std::numeric_limits<float>::infinity();
  constexpr bool operator<(const sphere_hit& o) const noexcept { return tmin < o.tmin; }
  constexpr bool operator<(float d) const noexcept { return tmin < d; }
};

// 64-bit LCG mapped to [0,1) floats: deterministic scene bytes.
struct lcg
{
  std::uint64_t s;
  float next()
  {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return float(s >> 40) * (1.0f / 16777216.0f);
  }
};

} // namespace

int
main(int argc, char** argv)
{
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  lcg g{ 12345 };
  const std::size_t n = 20000;
  std::vector<lbvh::triangle<float>> tris;
  tris.reserve(n);
  for (std::size_t i = 0; i < n; i++) {
    lbvh::vec3<float> c{ g.next(), g.next(), g.next() };
    lbvh::vec3<float> v[3];
    for (auto& p : v)
      p = { c.x + (g.next() - 0.5f) * 0.02f, c.y + (g.next() - 0.5f) * 0.02f,
            c.z + (g.next() - 0.5f) * 0.02f };
    tris.emplace_back(v[0], v[1], v[2]);
  }

  lbvh::triangle_aabb_converter<float> cvt;
  lbvh::builder<float> builder;
  auto bvh = builder(tris.data(), tris.size(), cvt);

  // the reference example's pattern: user primitive + lambda converter
  sphere spheres[3]{ { 1, { 5, 1, -5 } }, { 1, { 0, 2, -5 } }, { 1, { -5, 1, -5 } } };
  auto sphere_to_box = [](sphere s) -> lbvh::aabb<float> {
    return lbvh::aabb<float>{ { s.pos[0] - s.radius, s.pos[1] - s.radius, s.pos[2] - s.radius },
                              { s.pos[0] + s.radius, s.pos[1] + s.radius, s.pos[2] + s.radius } };
  };
  auto sbvh = builder(spheres, 3, sphere_to_box);
  lbvh::traverser<float, sphere, sphere_hit> straverser(sbvh, spheres);
  auto isect_sphere = [](const sphere& s, const lbvh::ray<float>& r) -> sphere_hit {
    using namespace lbvh::math;
    lbvh::vec3<float> c{ s.pos[0], s.pos[1], s.pos[2] };
    auto d = r.pos - c;
    float a = dot(r.dir, r.dir), b = 2 * dot(r.dir, d), cc = dot(d, d) - s.radius * s.radius;
    float disc = b * b - 4 * a * cc;
    if (disc < 0)
      return sphere_hit{};
    float t = (-b - std::sqrt(disc)) / (2 * a);
    return sphere_hit{ t };
  };
  lbvh::ray<float> sr{ { 0, 0, 5 }, { 0, 0.2f, -1 } }; // aims at the centre of sphere 1
  auto sh = straverser(sr, isect_sphere);
  if (!sh || sh.primitive != 1) {
    std::fprintf(stderr, "sphere scene: expected to hit sphere 1, got %u\n", unsigned(sh.primitive));
    return 1;
  }

  // closest hit: host traverser over the downloaded tree vs the device tree
  const std::size_t m = 4096;
  std::vector<lbvh::ray<float>> rays(m);
  for (std::size_t i = 0; i < m; i++) {
    float x = float(i % 64) / 63.0f, y = float(i / 64) / 63.0f;
    rays[i].pos = { 0.5f, 0.5f, -1.0f };
    rays[i].dir = lbvh::math::normalize(lbvh::vec3<float>{ x - 0.5f, y - 0.5f, 1.0f });
  }
  lbvh::traverser<float, lbvh::triangle<float>, lbvh::triangle_intersection<float>> tr(
    bvh, tris.data());
  lbvh::triangle_intersector<float> isect;
  std::vector<std::uint32_t> hp(m);
  std::vector<float> ht(m);
  for (std::size_t i = 0; i < m; i++) {
    auto h = tr(rays[i], isect);
    hp[i] = h.primitive;
    ht[i] = h.info.distance;
  }
  lbvh::gpu::device_tree dt(tris.data(), tris.size());
  std::vector<std::uint32_t> dp(m);
  std::vector<float> dtt(m);
  dt.trace(rays.data(), m, dp.data(), dtt.data());
  for (std::size_t i = 0; i < m; i++) {
    if (dp[i] != hp[i] || std::memcmp(&dtt[i], &ht[i], 4) != 0) {
      std::fprintf(stderr, "ray %zu: device (%u, %g) != host (%u, %g)\n", i, dp[i], dtt[i], hp[i],
                   ht[i]);
      return 1;
    }
  }
  auto back = dt.download();
  if (std::memcmp(back.data(), bvh.data(), (n - 1) * 32) != 0) {
    std::fprintf(stderr, "device_tree nodes differ from builder nodes\n");
    return 1;
  }

  // builder<double>: FP64 triangles of the same vertices, u64 codes and
  // 64-byte nodes; the host traverser runs over the double tree unchanged
  std::vector<lbvh::triangle<double>> tris64;
  tris64.reserve(n);
  for (std::size_t i = 0; i < n; i++) {
    lcg h{ 777 + i };
    lbvh::vec3<double> c{ h.next(), h.next(), h.next() };
    lbvh::vec3<double> v[3];
    for (auto& p : v)
      p = { c.x + (h.next() - 0.5) * 0.02, c.y + (h.next() - 0.5) * 0.02,
            c.z + (h.next() - 0.5) * 0.02 };
    tris64.emplace_back(v[0], v[1], v[2]);
  }
  lbvh::builder<double> builder64;
  auto bvh64 = builder64(tris64.data(), tris64.size(), lbvh::triangle_aabb_converter<double>());
  if (bvh64.size() != n - 1) {
    std::fprintf(stderr, "builder<double>: %zu nodes\n", std::size_t(bvh64.size()));
    return 1;
  }
  lbvh::traverser<double, lbvh::triangle<double>, lbvh::triangle_intersection<double>> tr64(
    bvh64, tris64.data());
  lbvh::triangle_intersector<double> isect64;
  std::size_t hits64 = 0;
  std::vector<lbvh::ray<double>> rays64(256);
  for (std::size_t i = 0; i < 256; i++) {
    rays64[i].pos = { 0.5, 0.5, -1.0 };
    rays64[i].dir = lbvh::math::normalize(
      lbvh::vec3<double>{ double(i % 16) / 15.0 - 0.5, double(i / 16) / 15.0 - 0.5, 1.0 });
  }
  // the same rays traced on the device in one batch
  std::vector<std::uint64_t> gp(256);
  std::vector<double> gt(256);
  std::vector<lbvh::vec2<double>> guv(256);
  lbvh::gpu::trace_closest(bvh64, tris64.data(), tris64.size(), rays64.data(), rays64.size(),
                           gp.data(), gt.data(), guv.data());
  for (std::size_t i = 0; i < 256; i++) {
    const lbvh::ray<double>& r = rays64[i];
    auto h = tr64(r, isect64);
    if (gp[i] != h.primitive || std::memcmp(&gt[i], &h.info.distance, 8) != 0 ||
        (h && std::memcmp(&guv[i], &h.info.uv, 16) != 0)) {
      std::fprintf(stderr, "builder<double> ray %zu: device hit %llu at %g, host %llu at %g\n", i,
                   (unsigned long long)gp[i], gt[i], (unsigned long long)h.primitive,
                   h.info.distance);
      return 1;
    }
    // brute force over every triangle, first hit wins ties as in the traverser
    lbvh::triangle_intersection<double> best;
    std::uint64_t bi = ~std::uint64_t(0);
    for (std::size_t k = 0; k < n; k++) {
      auto info = isect64(tris64[k], r);
      if (info < best) {
        best = info;
        bi = k;
      }
    }
    if ((h ? h.info.distance : best.distance) != best.distance || (!h && bi != ~std::uint64_t(0))) {
      std::fprintf(stderr, "builder<double> ray %zu: tree %g, brute force %g\n", i,
                   h ? h.info.distance : -1.0, best.distance);
      return 1;
    }
    hits64 += h ? 1 : 0;
  }

  // the same FP64 scene with primitive-id leaves through the session options
  lbvh::gpu::default_options().leaf_order = LBVH_LEAF_PRIMITIVE_ID;
  auto bvh64_pid = builder64(tris64.data(), tris64.size(), lbvh::triangle_aabb_converter<double>());
  lbvh::gpu::default_options().leaf_order = LBVH_LEAF_REFERENCE;

  FILE* f = std::fopen(argv[1], "wb");
  if (!f)
    return 2;
  std::uint64_t nn = n, mm = m;
  std::fwrite(&nn, 8, 1, f);
  std::fwrite(tris.data(), 48, n, f);
  std::fwrite(bvh.data(), 32, n - 1, f);
  std::fwrite(&mm, 8, 1, f);
  std::fwrite(rays.data(), 32, m, f);
  std::fwrite(hp.data(), 4, m, f);
  std::fwrite(ht.data(), 4, m, f);
  std::fwrite(sbvh.data(), 32, 2, f);
  std::fwrite(&nn, 8, 1, f);
  std::fwrite(tris64.data(), 96, n, f);
  std::fwrite(bvh64.data(), 64, n - 1, f);
  std::fwrite(bvh64_pid.data(), 64, n - 1, f);
  std::fclose(f);
  std::printf("dropin ok: %zu triangles, %zu rays, %zu hits; builder<double>: %zu of 256 hits\n",
              n, m,
              std::size_t(std::count_if(hp.begin(), hp.end(),
                                        [](std::uint32_t p) { return p != 0xffffffffu; })),
              hits64);
  return 0;
}
