/*
 * Synthetic scene and ray generators for the LBVH workloads.
 *
 * The workloads are fixed by seed and shape; the generator itself is pinned
 * here so that every consumer (the HIP path, the CPU oracle, the reference
 * driver) sees byte-identical inputs:
 *
 *   - PRNG: xoshiro256** seeded through SplitMix64.
 *   - uniform float on [a,b): a + (b - a) * u with u = (x >> 40) * 2^-24
 *     (exactly representable), evaluated in FP32 with contraction off.
 *   - Gaussian: Box-Muller in double, rounded to float by the caller.
 *
 * Triangles are produced as vertex triples (9 floats) and turned into the
 * 48-byte lbvh::triangle<float> record {p0, e0 = a - b, e1 = c - a,
 * normal = cross(e0, e1)} exactly as its constructor does
 * (lbvh.h:1967-1972, cross at lbvh.h:609-616).
 *
 * Build with -ffp-contract=off: the normal and the uniform mapping are
 * contraction-sensitive.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__FP_FAST_FMA) && !defined(WL_ALLOW_FMA)
/* not an error: FMA availability is fine as long as contraction is off */
#endif

typedef struct
{
  uint64_t s[4];
} wl_rng;

static uint64_t
wl_splitmix64(uint64_t* x)
{
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static void
wl_seed(wl_rng* r, uint64_t seed)
{
  uint64_t x = seed;
  for (int i = 0; i < 4; i++)
    r->s[i] = wl_splitmix64(&x);
}

static inline uint64_t
wl_rotl(uint64_t x, int k)
{
  return (x << k) | (x >> (64 - k));
}

static inline uint64_t
wl_next(wl_rng* r)
{
  uint64_t* s = r->s;
  const uint64_t result = wl_rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = wl_rotl(s[3], 45);
  return result;
}

static inline float
wl_u24(wl_rng* r)
{
  return (float)(wl_next(r) >> 40) * 0x1p-24f;
}

static inline float
wl_uniform(wl_rng* r, float a, float b)
{
  float u = wl_u24(r);
  float w = b - a;
  float p = w * u;
  return a + p;
}

static inline double
wl_u53(wl_rng* r)
{
  return (double)(wl_next(r) >> 11) * 0x1p-53;
}

/* One standard normal pair by Box-Muller. */
static void
wl_gauss2(wl_rng* r, double* z0, double* z1)
{
  double u1 = wl_u53(r);
  double u2 = wl_u53(r);
  double rad = sqrt(-2.0 * log(1.0 - u1));
  double th = 6.283185307179586 * u2;
  *z0 = rad * cos(th);
  *z1 = rad * sin(th);
}

/* lbvh::triangle<float>(a, b, c): p0 = a, e0 = a - b, e1 = c - a,
 * normal = cross(e0, e1). */
void
wl_make_triangles(const float* verts, uint64_t n, float* tri)
{
  for (uint64_t i = 0; i < n; i++) {
    const float* v = verts + 9 * i;
    float* t = tri + 12 * i;
    float e0[3], e1[3];
    for (int k = 0; k < 3; k++) {
      e0[k] = v[k] - v[3 + k];
      e1[k] = v[6 + k] - v[k];
    }
    float ay_bz = e0[1] * e1[2], az_by = e0[2] * e1[1];
    float az_bx = e0[2] * e1[0], ax_bz = e0[0] * e1[2];
    float ax_by = e0[0] * e1[1], ay_bx = e0[1] * e1[0];
    t[0] = v[0];
    t[1] = v[1];
    t[2] = v[2];
    t[3] = e0[0];
    t[4] = e0[1];
    t[5] = e0[2];
    t[6] = e1[0];
    t[7] = e1[1];
    t[8] = e1[2];
    t[9] = ay_bz - az_by;
    t[10] = az_bx - ax_bz;
    t[11] = ax_by - ay_bx;
  }
}

/* Centres uniform in [lo, lo+1)^3, each vertex = centre + U[-h, h)^3.
 * Draw order per triangle: cx cy cz, then v0 xyz, v1 xyz, v2 xyz. */
void
wl_uniform_scene(uint64_t seed, uint64_t n, float h, const float* lo,
                 float* verts)
{
  wl_rng r;
  wl_seed(&r, seed);
  float base[3] = { 0.0f, 0.0f, 0.0f };
  if (lo)
    memcpy(base, lo, sizeof(base));
  for (uint64_t i = 0; i < n; i++) {
    float c[3];
    for (int k = 0; k < 3; k++)
      c[k] = wl_uniform(&r, base[k], base[k] + 1.0f);
    for (int v = 0; v < 3; v++)
      for (int k = 0; k < 3; k++)
        verts[9 * i + 3 * v + k] = c[k] + wl_uniform(&r, -h, h);
  }
}

/* Centres from `nclusters` isotropic Gaussians (means uniform in [0,1)^3,
 * standard deviation sigma, equal weights), clamped to [0,1]; each vertex =
 * centre + U[-h, h)^3.  Cluster choice uses the top 6+ bits of one draw. */
void
wl_cluster_scene(uint64_t seed, uint64_t n, uint32_t nclusters, float sigma,
                 float h, float* verts)
{
  wl_rng r;
  wl_seed(&r, seed);
  float means[3 * 1024];
  if (nclusters == 0 || nclusters > 1024)
    return;
  for (uint32_t c = 0; c < nclusters; c++)
    for (int k = 0; k < 3; k++)
      means[3 * c + k] = wl_u24(&r);
  for (uint64_t i = 0; i < n; i++) {
    uint32_t ci = (uint32_t)(((wl_next(&r) >> 32) * (uint64_t)nclusters) >> 32);
    double z[4];
    wl_gauss2(&r, &z[0], &z[1]);
    wl_gauss2(&r, &z[2], &z[3]);
    float c[3];
    for (int k = 0; k < 3; k++) {
      double v = (double)means[3 * ci + k] + (double)sigma * z[k];
      float f = (float)v;
      f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
      c[k] = f;
    }
    for (int v = 0; v < 3; v++)
      for (int k = 0; k < 3; k++)
        verts[9 * i + 3 * v + k] = c[k] + wl_uniform(&r, -h, h);
  }
}

/* Adds U[-amp, amp)^3 to every vertex (topology-preserving perturbation). */
void
wl_perturb(uint64_t seed, float* verts, uint64_t n, float amp)
{
  wl_rng r;
  wl_seed(&r, seed);
  for (uint64_t i = 0; i < 9 * n; i++)
    verts[i] = verts[i] + wl_uniform(&r, -amp, amp);
}

static void
wl_normalize(float v[3])
{
  /* lbvh::math::normalize (lbvh.h:725-734): v * (1 / sqrt(dot(v, v))) */
  float d = 0.0f;
  float xx = v[0] * v[0], yy = v[1] * v[1], zz = v[2] * v[2];
  d = d + xx;
  d = d + yy;
  d = d + zz;
  float l_inv = 1.0f / sqrtf(d);
  v[0] = v[0] * l_inv;
  v[1] = v[1] * l_inv;
  v[2] = v[2] * l_inv;
}

/* Pinhole primary rays, row-major (y outer), as lbvh::ray<float>
 * {pos, dir, tmin, tmax}.  dir = normalize(((2(x+1/2)/W - 1) * tanh * aspect,
 * (1 - 2(y+1/2)/H) * tanh, 1)), looking +z. */
void
wl_primary_rays(uint32_t w, uint32_t h, const float* eye, float tan_half,
                float tmin, float tmax, float* rays)
{
  float aspect = (float)w / (float)h;
  for (uint32_t y = 0; y < h; y++) {
    for (uint32_t x = 0; x < w; x++) {
      float* r = rays + 8 * ((uint64_t)y * w + x);
      float sx = (2.0f * ((float)x + 0.5f)) / (float)w - 1.0f;
      float sy = 1.0f - (2.0f * ((float)y + 0.5f)) / (float)h;
      float d[3] = { (sx * tan_half) * aspect, sy * tan_half, 1.0f };
      wl_normalize(d);
      r[0] = eye[0];
      r[1] = eye[1];
      r[2] = eye[2];
      r[3] = d[0];
      r[4] = d[1];
      r[5] = d[2];
      r[6] = tmin;
      r[7] = tmax;
    }
  }
}

/* One cosine-weighted diffuse bounce per hit.  Origin = o + t*d; the
 * geometric normal (the triangle's stored normal, normalized) is flipped to
 * face the incoming ray; the direction is sampled about it.  Writes the
 * bounce rays compacted in primary-ray order and returns their count.  When
 * `src` is non-null, src[k] receives the primary-ray index of bounce k. */
uint64_t
wl_bounce_rays(uint64_t seed, const float* tri, const float* rays,
               uint64_t nrays, const uint32_t* hit_prim, const float* hit_t,
               float tmin, float tmax, float* out, uint32_t* src)
{
  wl_rng r;
  wl_seed(&r, seed);
  uint64_t m = 0;
  for (uint64_t i = 0; i < nrays; i++) {
    if (hit_prim[i] == 0xffffffffu)
      continue;
    const float* ry = rays + 8 * i;
    const float* tr = tri + 12 * (uint64_t)hit_prim[i];
    double o[3], d[3], nrm[3];
    for (int k = 0; k < 3; k++) {
      o[k] = ry[k];
      d[k] = ry[3 + k];
      nrm[k] = tr[9 + k];
    }
    double nl = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    if (!(nl > 0.0)) {
      nrm[0] = 0.0;
      nrm[1] = 0.0;
      nrm[2] = 1.0;
      nl = 1.0;
    }
    for (int k = 0; k < 3; k++)
      nrm[k] /= nl;
    if (nrm[0] * d[0] + nrm[1] * d[1] + nrm[2] * d[2] > 0.0)
      for (int k = 0; k < 3; k++)
        nrm[k] = -nrm[k];
    /* orthonormal basis (Frisvad-style with the branch-free Duff variant) */
    double s = nrm[2] >= 0.0 ? 1.0 : -1.0;
    double a = -1.0 / (s + nrm[2]);
    double b = nrm[0] * nrm[1] * a;
    double t1[3] = { 1.0 + s * nrm[0] * nrm[0] * a, s * b, -s * nrm[0] };
    double t2[3] = { b, s + nrm[1] * nrm[1] * a, -nrm[1] };
    double r1 = wl_u53(&r), r2 = wl_u53(&r);
    double phi = 6.283185307179586 * r1;
    double rad = sqrt(r2);
    double lx = rad * cos(phi), ly = rad * sin(phi), lz = sqrt(1.0 - r2);
    float* q = out + 8 * m;
    float t = hit_t[i];
    for (int k = 0; k < 3; k++) {
      float dk = ry[3 + k];
      float sk = t * dk;
      q[k] = ry[k] + sk;
      q[3 + k] = (float)(lx * t1[k] + ly * t2[k] + lz * nrm[k]);
    }
    q[6] = tmin;
    q[7] = tmax;
    if (src)
      src[m] = (uint32_t)i;
    m++;
  }
  return m;
}

/* Rays with origins uniform on the surface of the box [lo, hi] (face chosen
 * in proportion to its area) and directions uniform on the unit sphere. */
void
wl_box_rays(uint64_t seed, const float* lo, const float* hi, uint64_t nrays,
            float tmin, float tmax, float* rays)
{
  wl_rng r;
  wl_seed(&r, seed);
  double ext[3];
  for (int k = 0; k < 3; k++)
    ext[k] = (double)hi[k] - (double)lo[k];
  double area[3] = { ext[1] * ext[2], ext[0] * ext[2], ext[0] * ext[1] };
  double tot = 2.0 * (area[0] + area[1] + area[2]);
  for (uint64_t i = 0; i < nrays; i++) {
    float* q = rays + 8 * i;
    double pick = wl_u53(&r) * tot;
    int face = 0;
    double acc = 0.0;
    for (face = 0; face < 5; face++) {
      acc += area[face >> 1];
      if (pick < acc)
        break;
    }
    int ax = face >> 1;
    double u = wl_u53(&r), v = wl_u53(&r);
    int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    double p[3];
    p[ax] = (face & 1) ? hi[ax] : lo[ax];
    p[a1] = lo[a1] + u * ext[a1];
    p[a2] = lo[a2] + v * ext[a2];
    double z = 2.0 * wl_u53(&r) - 1.0;
    double phi = 6.283185307179586 * wl_u53(&r);
    double rr = sqrt(1.0 - z * z);
    double d[3] = { rr * cos(phi), rr * sin(phi), z };
    for (int k = 0; k < 3; k++) {
      q[k] = (float)p[k];
      q[3 + k] = (float)d[k];
    }
    q[6] = tmin;
    q[7] = tmax;
  }
}

/* Scene sizes log-uniform in [lo, hi]: round(exp(U[ln lo, ln hi])). */
void
wl_loguniform_sizes(uint64_t seed, uint64_t count, double lo, double hi,
                    uint64_t* out)
{
  wl_rng r;
  wl_seed(&r, seed);
  double a = log(lo), b = log(hi);
  for (uint64_t i = 0; i < count; i++) {
    double x = exp(a + (b - a) * wl_u53(&r));
    uint64_t v = (uint64_t)floor(x + 0.5);
    if (v < (uint64_t)lo)
      v = (uint64_t)lo;
    if (v > (uint64_t)hi)
      v = (uint64_t)hi;
    out[i] = v;
  }
}

/* Raw generator access for request-stream synthesis. */
void
wl_uniform_u53(uint64_t seed, uint64_t count, double* out)
{
  wl_rng r;
  wl_seed(&r, seed);
  for (uint64_t i = 0; i < count; i++)
    out[i] = wl_u53(&r);
}

/* Bounds of the vertices a, b = a - e0, c = a + e1 of n triangles (48-byte
 * lbvh::triangle<float> records), in float arithmetic. */
void
wl_tri_bounds(const float* tri, uint64_t n, float* lo, float* hi)
{
  for (int k = 0; k < 3; k++) {
    lo[k] = INFINITY;
    hi[k] = -INFINITY;
  }
  for (uint64_t i = 0; i < n; i++) {
    const float* t = tri + 12 * i;
    for (int k = 0; k < 3; k++) {
      float a = t[k];
      float b = a - t[3 + k];
      float c = a + t[6 + k];
      float mn = a < b ? a : b;
      float mx = a < b ? b : a;
      mn = c < mn ? c : mn;
      mx = c > mx ? c : mx;
      lo[k] = mn < lo[k] ? mn : lo[k];
      hi[k] = mx > hi[k] ? mx : hi[k];
    }
  }
}
