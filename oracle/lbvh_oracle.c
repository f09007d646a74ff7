/*
 * ORACLE — test infrastructure only.  CPU restatement of the tay10r/lbvh
 * build and closest-hit path; see lbvh_oracle.h for scope and pinning.
 * Compile with -ffp-contract=off, no fast-math.
 */
#include "lbvh_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- scalar semantics (lbvh.h:566-580, libstdc++ std::min/std::max) ----- */

/* std::min(a, b) == (b < a) ? b : a : keeps `a` on ties and on NaN in b. */
static inline float
std_min(float a, float b)
{
  return (b < a) ? b : a;
}
/* std::max(a, b) == (a < b) ? b : a */
static inline float
std_max(float a, float b)
{
  return (a < b) ? b : a;
}
/* lbvh::math::min (lbvh.h:567-572) */
static inline float
m_min(float a, float b)
{
  return (a < b) ? a : b;
}
/* lbvh::math::max (lbvh.h:575-580) */
static inline float
m_max(float a, float b)
{
  return (a > b) ? a : b;
}

/* lbvh::math::dot (lbvh.h:621-630): accumulates from 0 in x, y, z order. */
static inline float
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

/* lbvh::math::cross (lbvh.h:609-616) */
static inline void
cross3(const float* a, const float* b, float* o)
{
  float ay_bz = a[1] * b[2], az_by = a[2] * b[1];
  float az_bx = a[2] * b[0], ax_bz = a[0] * b[2];
  float ax_by = a[0] * b[1], ay_bx = a[1] * b[0];
  o[0] = ay_bz - az_by;
  o[1] = az_bx - ax_bz;
  o[2] = ax_by - ay_bx;
}

/* ---- A2: triangle_aabb_converter (lbvh.h:2001-2012) ---------------------- */

void
orc_triangle_aabb(const float* tri, orc_aabb* out)
{
  const float* p0 = tri;
  const float* e0 = tri + 3;
  const float* e1 = tri + 6;
  for (int k = 0; k < 3; k++) {
    float p1 = p0[k] - e0[k];
    float p2 = p0[k] + e1[k];
    float tmin = std_min(p0[k], p1);
    float tmax = std_max(p0[k], p1);
    out->min[k] = std_min(tmin, p2);
    out->max[k] = std_max(tmax, p2);
  }
}

/* center_of (lbvh.h:1160-1167): (min + max) * 0.5 */
static inline void
center_of(const orc_aabb* b, float* c)
{
  for (int k = 0; k < 3; k++) {
    float s = b->min[k] + b->max[k];
    c[k] = s * 0.5f;
  }
}

/* ---- A3: centroid_bounds_kernel + serial union (lbvh.h:1282-1293,
 *          1539-1543); empty box = +-inf (lbvh.h:1190-1200); union_of with
 *          std::min/std::max, accumulator first (lbvh.h:1208-1223). -------- */

void
orc_centroid_bounds(const float* tris, uint64_t n, orc_aabb* out)
{
  orc_aabb b;
  for (int k = 0; k < 3; k++) {
    b.min[k] = INFINITY;
    b.max[k] = -INFINITY;
  }
  for (uint64_t i = 0; i < n; i++) {
    orc_aabb box;
    float c[3];
    orc_triangle_aabb(tris + 12 * i, &box);
    center_of(&box, c);
    for (int k = 0; k < 3; k++) {
      b.min[k] = std_min(b.min[k], c[k]);
      b.max[k] = std_max(b.max[k], c[k]);
    }
  }
  *out = b;
}

/* ---- A4: morton_encoder<4>/<8> expansions (lbvh.h:1352-1358,
 *          1377-1385) and morton_curve_kernel quantisation (lbvh.h:1432-1473)
 */

static inline uint32_t
expand10(uint32_t n)
{
  n = (n | (n << 16)) & 0x030000ffu;
  n = (n | (n << 8)) & 0x0300f00fu;
  n = (n | (n << 4)) & 0x030c30c3u;
  n = (n | (n << 2)) & 0x09249249u;
  return n;
}

static inline uint64_t
expand20(uint64_t n)
{
  n = (n | n << 32) & 0x001f00000000ffffull;
  n = (n | n << 16) & 0x001f0000ff0000ffull;
  n = (n | n << 8) & 0x100f00f00f00f00full;
  n = (n | n << 4) & 0x10c30c30c30c30c3ull;
  n = (n | n << 2) & 0x1249249249249249ull;
  return n;
}

void
orc_morton_codes(const float* tris, uint64_t n, int code_bits,
                 const orc_aabb* bounds, void* codes)
{
  /* scale = D / (extent + 1), D = 1024 (u32 codes) or 2^20 (u64 codes) */
  float dom = (code_bits == 30) ? 1024.0f : 1048576.0f;
  float scale[3];
  for (int k = 0; k < 3; k++) {
    float ext = bounds->max[k] - bounds->min[k];
    float den = ext + 1.0f;
    scale[k] = dom / den;
  }
  for (uint64_t i = 0; i < n; i++) {
    orc_aabb box;
    float c[3];
    orc_triangle_aabb(tris + 12 * i, &box);
    center_of(&box, c);
    float q[3];
    for (int k = 0; k < 3; k++) {
      float d = c[k] - bounds->min[k];
      q[k] = d * scale[k];
    }
    if (code_bits == 30) {
      uint32_t x = (uint32_t)q[0], y = (uint32_t)q[1], z = (uint32_t)q[2];
      ((uint32_t*)codes)[i] =
        (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
    } else {
      uint64_t x = (uint64_t)q[0], y = (uint64_t)q[1], z = (uint64_t)q[2];
      ((uint64_t*)codes)[i] =
        (expand20(x) << 2) | (expand20(y) << 1) | expand20(z);
    }
  }
}

/* ---- A5: the reference sorts entries by code with an unstable std::sort
 *          (lbvh.h:1126-1134).  The engine defines sorted order as stable by
 *          (code, primitive id), one legal output of that sort; restated here
 *          as an LSD radix sort with 8-bit digits. -------------------------- */

void
orc_sort_pairs(int code_bits, const void* codes, uint64_t n,
               void* sorted_codes, uint32_t* ids)
{
  const int kb = (code_bits == 30) ? 4 : 8;
  uint64_t* ka = (uint64_t*)malloc(n * sizeof(uint64_t));
  uint64_t* kt = (uint64_t*)malloc(n * sizeof(uint64_t));
  uint32_t* va = (uint32_t*)malloc(n * sizeof(uint32_t));
  uint32_t* vt = (uint32_t*)malloc(n * sizeof(uint32_t));
  for (uint64_t i = 0; i < n; i++) {
    ka[i] = (code_bits == 30) ? ((const uint32_t*)codes)[i]
                              : ((const uint64_t*)codes)[i];
    va[i] = (uint32_t)i;
  }
  for (int pass = 0; pass < kb; pass++) {
    uint64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    int sh = 8 * pass;
    for (uint64_t i = 0; i < n; i++)
      cnt[((ka[i] >> sh) & 0xff) + 1]++;
    for (int d = 0; d < 256; d++)
      cnt[d + 1] += cnt[d];
    for (uint64_t i = 0; i < n; i++) {
      uint64_t pos = cnt[(ka[i] >> sh) & 0xff]++;
      kt[pos] = ka[i];
      vt[pos] = va[i];
    }
    uint64_t* tk = ka;
    ka = kt;
    kt = tk;
    uint32_t* tv = va;
    va = vt;
    vt = tv;
  }
  for (uint64_t i = 0; i < n; i++) {
    if (code_bits == 30)
      ((uint32_t*)sorted_codes)[i] = (uint32_t)ka[i];
    else
      ((uint64_t*)sorted_codes)[i] = ka[i];
    if (ids)
      ids[i] = va[i];
  }
  free(ka);
  free(kt);
  free(va);
  free(vt);
}

/* ---- A6: divide_node (lbvh.h:1588-1688) and builder_kernel
 *          (lbvh.h:1713-1731).  Offsets are int64 here; for n < 2^29 they
 *          give the same values as the reference's int32 offsets. ---------- */

typedef struct
{
  const void* codes;
  int64_t n;
  int wide; /* 0: u32 codes, 1: u64 codes */
} orc_table;

static inline int
clz32(uint32_t x)
{
  return x ? __builtin_clz(x) : 32;
}
static inline int
clz64(uint64_t x)
{
  return x ? __builtin_clzll(x) : 64;
}

/* calc_delta (lbvh.h:1602-1621) */
static inline int64_t
calc_delta(const orc_table* t, int64_t j, int64_t k)
{
  if (k < 0 || k >= t->n)
    return -1;
  if (!t->wide) {
    uint32_t a = ((const uint32_t*)t->codes)[j];
    uint32_t b = ((const uint32_t*)t->codes)[k];
    if (a == b)
      return clz32((uint32_t)(j ^ k)) + 32;
    return clz32(a ^ b);
  } else {
    uint64_t a = ((const uint64_t*)t->codes)[j];
    uint64_t b = ((const uint64_t*)t->codes)[k];
    if (a == b)
      return clz64((uint64_t)(j ^ k)) + 64;
    return clz64(a ^ b);
  }
}

static void
divide_node(const orc_table* t, int64_t i, int64_t* out_j, int64_t* out_gamma)
{
  int64_t dd = calc_delta(t, i, i + 1) - calc_delta(t, i, i - 1);
  int64_t d = (dd < 0) ? -1 : 1; /* sign (lbvh.h:941-946) */
  int64_t delta_min = calc_delta(t, i, i - d);
  int64_t l_max = 128;
  for (;;) {
    int64_t k = i + l_max * d;
    if (k >= t->n)
      break;
    if (!(calc_delta(t, i, k) > delta_min))
      break;
    l_max *= 4;
  }
  int64_t l = 0;
  for (int64_t div = 2;; div *= 2) {
    int64_t tt = l_max / div;
    int64_t k = i + (l + tt) * d;
    if (calc_delta(t, i, k) > delta_min)
      l += tt;
    if (tt == 1)
      break;
  }
  int64_t j = i + l * d;
  int64_t delta_node = calc_delta(t, i, j);
  int64_t s = 0;
  for (int64_t div = 2;; div *= 2) {
    int64_t tt = (l + (div - 1)) / div; /* ceil_div (lbvh.h:953-958) */
    int64_t k = i + (s + tt) * d;
    if (calc_delta(t, i, k) > delta_node)
      s += tt;
    if (div >= l)
      break;
  }
  *out_j = j;
  *out_gamma = i + s * d + (d < 0 ? d : 0);
}

int
orc_karras(int code_bits, const void* sorted_codes, uint64_t n,
           orc_node* nodes)
{
  if (n < 2)
    return -1;
  orc_table t = { sorted_codes, (int64_t)n, code_bits == 30 ? 0 : 1 };
  for (int64_t i = 0; i < (int64_t)n - 1; i++) {
    int64_t j, g;
    divide_node(&t, i, &j, &g);
    int64_t mn = (i < j) ? i : j;
    int64_t mx = (j > i) ? j : i;
    uint32_t lm = (mn == g) ? ORC_LEAF_BIT : 0u;
    uint32_t rm = (mx == g + 1) ? ORC_LEAF_BIT : 0u;
    nodes[i].left = (uint32_t)g | lm;
    nodes[i].right = (uint32_t)(g + 1) | rm;
  }
  return 0;
}

/* ---- A7/A9: builder::fit_boxes (lbvh.h:1827-1866).  A breadth-first list
 *          of internal nodes from the root, then boxes in reverse order:
 *          box = (left leaf ? cvt(prims[leaf]) : nodes[left].box), then
 *          union with the right child, left operand first.  `leaf_prim`
 *          maps a leaf index to the primitive it stands for; NULL means the
 *          identity, which is the reference's behaviour. ------------------- */

void
orc_fit(orc_node* nodes, uint64_t n, const float* tris,
        const uint32_t* leaf_prim)
{
  uint64_t m = n - 1;
  uint64_t* idx = (uint64_t*)malloc(m * sizeof(uint64_t));
  uint64_t cnt = 0;
  idx[cnt++] = 0;
  for (uint64_t i = 0; i < cnt; i++) {
    uint64_t j = idx[i];
    if (!(nodes[j].left & ORC_LEAF_BIT))
      idx[cnt++] = nodes[j].left;
    if (!(nodes[j].right & ORC_LEAF_BIT))
      idx[cnt++] = nodes[j].right;
  }
  for (uint64_t i = cnt; i > 0; i--) {
    orc_node* nd = &nodes[idx[i - 1]];
    orc_aabb a, b;
    if (nd->left & ORC_LEAF_BIT) {
      uint32_t li = nd->left & ~ORC_LEAF_BIT;
      uint32_t p = leaf_prim ? leaf_prim[li] : li;
      orc_triangle_aabb(tris + 12 * (uint64_t)p, &a);
    } else {
      a = nodes[nd->left].box;
    }
    if (nd->right & ORC_LEAF_BIT) {
      uint32_t ri = nd->right & ~ORC_LEAF_BIT;
      uint32_t p = leaf_prim ? leaf_prim[ri] : ri;
      orc_triangle_aabb(tris + 12 * (uint64_t)p, &b);
    } else {
      b = nodes[nd->right].box;
    }
    for (int k = 0; k < 3; k++) {
      nd->box.min[k] = std_min(a.min[k], b.min[k]);
      nd->box.max[k] = std_max(a.max[k], b.max[k]);
    }
  }
  free(idx);
}

/* builder::operator() (lbvh.h:1796-1822) with the engine's two leaf orders. */
int
orc_build(const float* tris, uint64_t n, int code_bits, int leaf_order,
          orc_node* nodes, orc_aabb* bounds_out, void* sorted_codes_out,
          uint32_t* ids_out)
{
  if (n < 2 || (code_bits != 30 && code_bits != 63))
    return -1;
  size_t cw = (code_bits == 30) ? 4 : 8;
  orc_aabb bounds;
  orc_centroid_bounds(tris, n, &bounds);
  void* codes = malloc(n * cw);
  void* sorted = malloc(n * cw);
  uint32_t* ids = (uint32_t*)malloc(n * sizeof(uint32_t));
  orc_morton_codes(tris, n, code_bits, &bounds, codes);
  orc_sort_pairs(code_bits, codes, n, sorted, ids);
  memset(nodes, 0, (n - 1) * sizeof(orc_node));
  orc_karras(code_bits, sorted, n, nodes);
  if (leaf_order == ORC_LEAF_PRIMITIVE_ID) {
    for (uint64_t i = 0; i < n - 1; i++) {
      if (nodes[i].left & ORC_LEAF_BIT)
        nodes[i].left = ids[nodes[i].left & ~ORC_LEAF_BIT] | ORC_LEAF_BIT;
      if (nodes[i].right & ORC_LEAF_BIT)
        nodes[i].right = ids[nodes[i].right & ~ORC_LEAF_BIT] | ORC_LEAF_BIT;
    }
  }
  /* leaf indices already name primitives in both orders */
  orc_fit(nodes, n, tris, NULL);
  if (bounds_out)
    *bounds_out = bounds;
  if (sorted_codes_out)
    memcpy(sorted_codes_out, sorted, n * cw);
  if (ids_out)
    memcpy(ids_out, ids, n * sizeof(uint32_t));
  free(codes);
  free(sorted);
  free(ids);
  return 0;
}

/* ---- builder<double> (lbvh.h:549-558, 1796-1822 with scalar_type = double):
 *      u64 codes (morton_domain<8>, morton_encoder<8>), u64 node indices with
 *      the leaf flag in bit 63, FP64 boxes.  Input is the converter's boxes
 *      (boxes48[i] = converter(primitives[i])). -------------------------- */

static inline double
std_min_d(double a, double b)
{
  return (b < a) ? b : a;
}
static inline double
std_max_d(double a, double b)
{
  return (a < b) ? b : a;
}

/* triangle_aabb_converter<double> (lbvh.h:2001-2012) over triangle<double>
 * records of 12 doubles {p0, e0, e1, normal}. */
void
orc_triangle_aabb_f64(const double* tri96, uint64_t n, double* boxes48)
{
  for (uint64_t i = 0; i < n; i++) {
    const double* t = tri96 + 12 * i;
    double* o = boxes48 + 6 * i;
    for (int k = 0; k < 3; k++) {
      double p0 = t[k], p1 = t[k] - t[3 + k], p2 = t[k] + t[6 + k];
      o[k] = std_min_d(std_min_d(p0, p1), p2);
      o[3 + k] = std_max_d(std_max_d(p0, p1), p2);
    }
  }
}

/* centroid_bounds_kernel (lbvh.h:1282-1293) and morton_curve_kernel
 * (lbvh.h:1432-1473) in FP64: centre = (min + max) * 0.5 (lbvh.h:1164-1167),
 * scale = 2^20 / (extent + 1), q = (c - min) * scale truncated to u64. */
void
orc_morton_codes_f64(const double* boxes48, uint64_t n, double* bounds6,
                     uint64_t* codes)
{
  double lo[3] = { INFINITY, INFINITY, INFINITY };
  double hi[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (uint64_t i = 0; i < n; i++) {
    const double* b = boxes48 + 6 * i;
    for (int k = 0; k < 3; k++) {
      double c = (b[k] + b[3 + k]) * 0.5;
      lo[k] = std_min_d(lo[k], c);
      hi[k] = std_max_d(hi[k], c);
    }
  }
  double scale[3];
  for (int k = 0; k < 3; k++)
    scale[k] = 1048576.0 / ((hi[k] - lo[k]) + 1.0);
  for (uint64_t i = 0; i < n; i++) {
    const double* b = boxes48 + 6 * i;
    uint64_t q[3];
    for (int k = 0; k < 3; k++) {
      double c = (b[k] + b[3 + k]) * 0.5;
      q[k] = (uint64_t)((c - lo[k]) * scale[k]);
    }
    codes[i] = (expand20(q[0]) << 2) | (expand20(q[1]) << 1) | expand20(q[2]);
  }
  if (bounds6) {
    for (int k = 0; k < 3; k++) {
      bounds6[k] = lo[k];
      bounds6[3 + k] = hi[k];
    }
  }
}

/* builder_kernel<uint64_t, double> (lbvh.h:1713-1731) over sorted u64 codes. */
int
orc_karras_f64(const uint64_t* sorted_codes, uint64_t n, orc_node64* nodes)
{
  if (n < 2)
    return -1;
  orc_table t = { sorted_codes, (int64_t)n, 1 };
  for (int64_t i = 0; i < (int64_t)n - 1; i++) {
    int64_t j, g;
    divide_node(&t, i, &j, &g);
    int64_t mn = (i < j) ? i : j;
    int64_t mx = (j > i) ? j : i;
    uint64_t lm = (mn == g) ? ORC_LEAF_BIT64 : 0u;
    uint64_t rm = (mx == g + 1) ? ORC_LEAF_BIT64 : 0u;
    nodes[i].left = (uint64_t)g | lm;
    nodes[i].right = (uint64_t)(g + 1) | rm;
  }
  return 0;
}

/* fit_boxes (lbvh.h:1827-1866) in FP64 over precomputed boxes. */
void
orc_fit_f64(orc_node64* nodes, uint64_t n, const double* boxes48)
{
  uint64_t m = n - 1;
  uint64_t* idx = (uint64_t*)malloc(m * sizeof(uint64_t));
  uint64_t cnt = 0;
  idx[cnt++] = 0;
  for (uint64_t i = 0; i < cnt; i++) {
    uint64_t j = idx[i];
    if (!(nodes[j].left & ORC_LEAF_BIT64))
      idx[cnt++] = nodes[j].left;
    if (!(nodes[j].right & ORC_LEAF_BIT64))
      idx[cnt++] = nodes[j].right;
  }
  for (uint64_t i = cnt; i > 0; i--) {
    orc_node64* nd = &nodes[idx[i - 1]];
    const double *a, *b;
    if (nd->left & ORC_LEAF_BIT64)
      a = boxes48 + 6 * (nd->left & ~ORC_LEAF_BIT64);
    else
      a = nodes[nd->left].min;
    if (nd->right & ORC_LEAF_BIT64)
      b = boxes48 + 6 * (nd->right & ~ORC_LEAF_BIT64);
    else
      b = nodes[nd->right].min;
    double box[6];
    for (int k = 0; k < 3; k++) {
      box[k] = std_min_d(a[k], b[k]);
      box[3 + k] = std_max_d(a[3 + k], b[3 + k]);
    }
    memcpy(nd->min, box, sizeof(box));
  }
  free(idx);
}

int
orc_build_f64(const double* boxes48, uint64_t n, int leaf_order,
              orc_node64* nodes, uint64_t* sorted_codes_out, uint32_t* ids_out)
{
  if (n < 2 || n >= (1ull << 32))
    return -1;
  uint64_t* codes = (uint64_t*)malloc(n * 8);
  uint64_t* sorted = (uint64_t*)malloc(n * 8);
  uint32_t* ids = (uint32_t*)malloc(n * sizeof(uint32_t));
  orc_morton_codes_f64(boxes48, n, NULL, codes);
  orc_sort_pairs(63, codes, n, sorted, ids);
  memset(nodes, 0, (n - 1) * sizeof(orc_node64));
  orc_karras_f64(sorted, n, nodes);
  if (leaf_order == ORC_LEAF_PRIMITIVE_ID) {
    for (uint64_t i = 0; i < n - 1; i++) {
      if (nodes[i].left & ORC_LEAF_BIT64)
        nodes[i].left = ids[nodes[i].left & ~ORC_LEAF_BIT64] | ORC_LEAF_BIT64;
      if (nodes[i].right & ORC_LEAF_BIT64)
        nodes[i].right = ids[nodes[i].right & ~ORC_LEAF_BIT64] | ORC_LEAF_BIT64;
    }
  }
  orc_fit_f64(nodes, n, boxes48);
  if (sorted_codes_out)
    memcpy(sorted_codes_out, sorted, n * 8);
  if (ids_out)
    memcpy(ids_out, ids, n * sizeof(uint32_t));
  free(codes);
  free(sorted);
  free(ids);
  return 0;
}

/* ---- A12: triangle_intersector (lbvh.h:2055-2075) ------------------------ */

static inline int
isect_reference(const float* tri, const float* ray, float* t_out, float* u_out,
                float* v_out)
{
  const float* p0 = tri;
  const float* e0 = tri + 3;
  const float* e1 = tri + 6;
  const float* nrm = tri + 9;
  const float* pos = ray;
  const float* dir = ray + 3;
  float c[3] = { p0[0] - pos[0], p0[1] - pos[1], p0[2] - pos[2] };
  float r[3];
  cross3(dir, c, r);
  float inv_det = 1.0f / dot3(nrm, dir);
  float u = dot3(r, e1) * inv_det;
  float v = dot3(r, e0) * inv_det;
  float w = (1.0f - u) - v;
  if ((u >= 0.0f) && (v >= 0.0f) && (w >= 0.0f)) {
    float t = dot3(nrm, c) * inv_det;
    if ((t >= ray[6]) && (t <= ray[7])) {
      *t_out = t;
      *u_out = u;
      *v_out = v;
      return 1;
    }
  }
  return 0;
}

/* Watertight ray/triangle test (Woop, Benthin, Wald, "Watertight
 * Ray/Triangle Intersection", JCGT 2(1), 2013), two-sided, on the vertices
 * the AABB converter reconstructs: A = p0, B = p0 - e0, C = p0 + e1.
 * u and v are the barycentric weights of B and C, the same meaning as the
 * reference intersector's u and v.  No FP64 fallback: a zero edge function
 * counts as inside. */
static inline int
isect_watertight(const float* tri, const float* ray, float* t_out,
                 float* u_out, float* v_out)
{
  const float* p0 = tri;
  const float* e0 = tri + 3;
  const float* e1 = tri + 6;
  const float* org = ray;
  const float* dir = ray + 3;
  float ax = fabsf(dir[0]), ay = fabsf(dir[1]), az = fabsf(dir[2]);
  int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
  int kx = kz + 1;
  if (kx == 3)
    kx = 0;
  int ky = kx + 1;
  if (ky == 3)
    ky = 0;
  if (dir[kz] < 0.0f) {
    int tmp = kx;
    kx = ky;
    ky = tmp;
  }
  float Sx = dir[kx] / dir[kz];
  float Sy = dir[ky] / dir[kz];
  float Sz = 1.0f / dir[kz];
  float va[3], vb[3], vc[3];
  for (int k = 0; k < 3; k++) {
    float b = p0[k] - e0[k];
    float c = p0[k] + e1[k];
    va[k] = p0[k] - org[k];
    vb[k] = b - org[k];
    vc[k] = c - org[k];
  }
  float Ax = va[kx] - Sx * va[kz], Ay = va[ky] - Sy * va[kz];
  float Bx = vb[kx] - Sx * vb[kz], By = vb[ky] - Sy * vb[kz];
  float Cx = vc[kx] - Sx * vc[kz], Cy = vc[ky] - Sy * vc[kz];
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
    return 0;
  float det = U + V + W;
  if (det == 0.0f)
    return 0;
  float Az = Sz * va[kz], Bz = Sz * vb[kz], Cz = Sz * vc[kz];
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t >= ray[6] && t <= ray[7]))
    return 0;
  /* weights: U for A, V for B, W for C */
  *t_out = t;
  *u_out = V * rcp;
  *v_out = W * rcp;
  return 1;
}

/* ---- A10: make_accel_ray (lbvh.h:991-1010) and the octant box test
 *          (lbvh.h:1071-1091), hit iff tmax >= max(0, tmin)
 *          (lbvh.h:1023-1026), math::min/max ternaries. ------------------- */

typedef struct
{
  float rcp[3];
  float inv_pos[3];
  int oct[3];
  int inv_oct[3];
} orc_accel;

static inline void
make_accel(const float* ray, orc_accel* a)
{
  for (int k = 0; k < 3; k++) {
    a->rcp[k] = 1.0f / ray[3 + k];
    float np = -ray[k];
    a->inv_pos[k] = np * a->rcp[k];
  }
  a->oct[0] = (ray[3] > 0.0f) ? 0 : 3;
  a->oct[1] = (ray[4] > 0.0f) ? 1 : 4;
  a->oct[2] = (ray[5] > 0.0f) ? 2 : 5;
  a->inv_oct[0] = 3 - a->oct[0];
  a->inv_oct[1] = 5 - a->oct[1];
  a->inv_oct[2] = 7 - a->oct[2];
}

static inline int
box_test(const orc_aabb* box, const orc_accel* a, float* tmin_out)
{
  const float b[6] = { box->min[0], box->min[1], box->min[2],
                       box->max[0], box->max[1], box->max[2] };
  float tn[3], tf[3];
  for (int k = 0; k < 3; k++) {
    float pn = b[a->oct[k]] * a->rcp[k];
    tn[k] = pn + a->inv_pos[k];
    float pf = b[a->inv_oct[k]] * a->rcp[k];
    tf[k] = pf + a->inv_pos[k];
  }
  float tmin = m_max(m_max(tn[0], tn[1]), tn[2]);
  float tmax = m_min(m_min(tf[0], tf[1]), tf[2]);
  *tmin_out = tmin;
  return tmax >= m_max(0.0f, tmin);
}

/* ---- A11: traverser::operator() (lbvh.h:1874-1942) with
 *          traversal_stack<float,128> (lbvh.h:1746-1789). ---------------- */

#define ORC_STACK 128

static inline int
better(int mode, float t, uint32_t id, float best_t, uint32_t best_id)
{
  if (mode == ORC_ISECT_REFERENCE)
    return t < best_t;
  return (t < best_t) || (t == best_t && id < best_id);
}

static void
trace_one(const orc_node* nodes, const float* tris, const float* ray, int mode,
          uint32_t* prim_out, float* t_out, float* uv_out, uint64_t* cnt)
{
  uint32_t st_node[ORC_STACK];
  float st_t[ORC_STACK];
  int sp = 0;
  st_node[sp] = 0;
  st_t[sp] = INFINITY;
  sp++;
  orc_accel acc;
  make_accel(ray, &acc);
  uint32_t best = 0xffffffffu;
  float best_t = INFINITY, best_u = 0.0f, best_v = 0.0f;
  uint64_t popped = 0, boxes = 0, tests = 0;
  while (sp > 0) {
    sp--;
    uint32_t ni = st_node[sp];
    float etmin = st_t[sp];
    if (best_t < etmin)
      continue;
    popped++;
    const orc_node* nd = &nodes[ni];
    float lt = 0.0f, rt = 0.0f;
    int lh = 0, rh = 0;
    for (int side = 0; side < 2; side++) {
      uint32_t ch = side ? nd->right : nd->left;
      if (ch & ORC_LEAF_BIT) {
        uint32_t p = ch & ~ORC_LEAF_BIT;
        float t, u, v;
        int hit;
        tests++;
        if (mode == ORC_ISECT_WATERTIGHT)
          hit = isect_watertight(tris + 12 * (uint64_t)p, ray, &t, &u, &v);
        else
          hit = isect_reference(tris + 12 * (uint64_t)p, ray, &t, &u, &v);
        if (hit && better(mode, t, p, best_t, best)) {
          best = p;
          best_t = t;
          best_u = u;
          best_v = v;
        }
      } else {
        float tm;
        boxes++;
        int h = box_test(&nodes[ch].box, &acc, &tm);
        if (side) {
          rh = h;
          rt = tm;
        } else {
          lh = h;
          lt = tm;
        }
      }
    }
    /* push far then near; a push onto a full stack is dropped */
#define PUSH(N, T)                                                             \
  do {                                                                         \
    if (sp < ORC_STACK) {                                                      \
      st_node[sp] = (N);                                                       \
      st_t[sp] = (T);                                                          \
      sp++;                                                                    \
    }                                                                          \
  } while (0)
    if (lh && rh) {
      if (lt < rt) {
        PUSH(nd->right, rt);
        PUSH(nd->left, lt);
      } else {
        PUSH(nd->left, lt);
        PUSH(nd->right, rt);
      }
    } else if (lh) {
      PUSH(nd->left, lt);
    } else if (rh) {
      PUSH(nd->right, rt);
    }
#undef PUSH
  }
  *prim_out = best;
  *t_out = best_t;
  if (uv_out) {
    uv_out[0] = best_u;
    uv_out[1] = best_v;
  }
  if (cnt) {
    cnt[0] += popped;
    cnt[1] += boxes;
    cnt[2] += tests;
  }
}

typedef struct
{
  const orc_node* nodes;
  const float* tris;
  uint64_t ntris;
  const float* rays;
  uint64_t nrays;
  int mode;
  uint32_t* prim;
  float* t;
  float* uv;
  uint64_t cnt[3];
  int tid, nth;
  int brute;
} orc_job;

static void
brute_one(const float* tris, uint64_t n, const float* ray, int mode,
          uint32_t* prim, float* t_out)
{
  uint32_t best = 0xffffffffu;
  float best_t = INFINITY;
  for (uint64_t i = 0; i < n; i++) {
    float t, u, v;
    int hit = (mode == ORC_ISECT_WATERTIGHT)
                ? isect_watertight(tris + 12 * i, ray, &t, &u, &v)
                : isect_reference(tris + 12 * i, ray, &t, &u, &v);
    if (hit && better(ORC_ISECT_REFERENCE_LOWID, t, (uint32_t)i, best_t, best)) {
      best = (uint32_t)i;
      best_t = t;
    }
  }
  *prim = best;
  *t_out = best_t;
}

static void*
job_main(void* arg)
{
  orc_job* j = (orc_job*)arg;
  j->cnt[0] = j->cnt[1] = j->cnt[2] = 0;
  /* interleaved assignment, as the reference harness's rows
   * (lbvh_test.cpp:154) */
  for (uint64_t i = (uint64_t)j->tid; i < j->nrays; i += (uint64_t)j->nth) {
    if (j->brute)
      brute_one(j->tris, j->ntris, j->rays + 8 * i, j->mode, &j->prim[i],
                &j->t[i]);
    else
      trace_one(j->nodes, j->tris, j->rays + 8 * i, j->mode, &j->prim[i],
                &j->t[i], j->uv ? j->uv + 2 * i : NULL, j->cnt);
  }
  return NULL;
}

static void
run_jobs(orc_job* proto, int nthreads)
{
  if (nthreads < 1)
    nthreads = 1;
  if (nthreads > 256)
    nthreads = 256;
  orc_job* jobs = (orc_job*)calloc((size_t)nthreads, sizeof(orc_job));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int i = 0; i < nthreads; i++) {
    jobs[i] = *proto;
    jobs[i].tid = i;
    jobs[i].nth = nthreads;
  }
  for (int i = 1; i < nthreads; i++)
    pthread_create(&th[i], NULL, job_main, &jobs[i]);
  job_main(&jobs[0]);
  for (int i = 1; i < nthreads; i++)
    pthread_join(th[i], NULL);
  proto->cnt[0] = proto->cnt[1] = proto->cnt[2] = 0;
  for (int i = 0; i < nthreads; i++)
    for (int k = 0; k < 3; k++)
      proto->cnt[k] += jobs[i].cnt[k];
  free(jobs);
  free(th);
}

void
orc_trace(const orc_node* nodes, uint64_t n, const float* tris,
          const float* rays, uint64_t nrays, int mode, int nthreads,
          uint32_t* prim, float* t, float* uv, uint64_t* counters)
{
  orc_job j;
  memset(&j, 0, sizeof(j));
  j.nodes = nodes;
  j.tris = tris;
  j.ntris = n;
  j.rays = rays;
  j.nrays = nrays;
  j.mode = mode;
  j.prim = prim;
  j.t = t;
  j.uv = uv;
  run_jobs(&j, nthreads);
  if (counters)
    for (int k = 0; k < 3; k++)
      counters[k] = j.cnt[k];
}

/* Closest hit over all triangles, ties to the lowest id: a tree-independent
 * check of the traversal. */
void
orc_brute_closest(const float* tris, uint64_t n, const float* rays,
                  uint64_t nrays, int mode, int nthreads, uint32_t* prim,
                  float* t)
{
  orc_job j;
  memset(&j, 0, sizeof(j));
  j.tris = tris;
  j.ntris = n;
  j.rays = rays;
  j.nrays = nrays;
  j.mode = mode;
  j.prim = prim;
  j.t = t;
  j.brute = 1;
  run_jobs(&j, nthreads);
}

/* Ray/plane distance in FP64 for the plane through p0 spanned by e0, e1 of
 * the hit triangle (NaN for a miss): the error yardstick for reported t. */
void
orc_t_fp64(const float* tris, const float* rays, const uint32_t* prim,
           uint64_t nrays, double* t64)
{
  for (uint64_t i = 0; i < nrays; i++) {
    if (prim[i] == 0xffffffffu) {
      t64[i] = NAN;
      continue;
    }
    const float* tr = tris + 12 * (uint64_t)prim[i];
    const float* ry = rays + 8 * i;
    double e0[3] = { tr[3], tr[4], tr[5] }, e1[3] = { tr[6], tr[7], tr[8] };
    double nn[3] = { e0[1] * e1[2] - e0[2] * e1[1],
                     e0[2] * e1[0] - e0[0] * e1[2],
                     e0[0] * e1[1] - e0[1] * e1[0] };
    double c[3] = { (double)tr[0] - ry[0], (double)tr[1] - ry[1],
                    (double)tr[2] - ry[2] };
    double num = nn[0] * c[0] + nn[1] * c[1] + nn[2] * c[2];
    double den = nn[0] * ry[3] + nn[1] * ry[4] + nn[2] * ry[5];
    t64[i] = num / den;
  }
}

/* ---- traverser<double> (lbvh.h:1874-1942 with scalar_type = double) over
 *      node<double> records, with triangle_intersector<double>
 *      (lbvh.h:2055-2075), make_accel_ray (lbvh.h:991-1010) and the octant
 *      box test (lbvh.h:1071-1091) in FP64.  A leaf value names the
 *      triangle<double> record it tests (left_leaf_index, lbvh.h:1904-1905,
 *      1915-1916): the sorted position for reference leaves, the primitive id
 *      for primitive-id leaves.  Hits report that value; a miss is
 *      UINT64_MAX (intersection_container<double>, lbvh.h:364-366). ------ */

static inline double
dm_min(double a, double b)
{
  return (a < b) ? a : b;
}

static inline double
dm_max(double a, double b)
{
  return (a > b) ? a : b;
}

/* lbvh::math::dot (lbvh.h:621-630): ((0 + x x') + y y') + z z' */
static inline double
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

/* lbvh::math::cross (lbvh.h:609-616) */
static inline void
dcross3(const double* a, const double* b, double* o)
{
  double ay_bz = a[1] * b[2], az_by = a[2] * b[1];
  double az_bx = a[2] * b[0], ax_bz = a[0] * b[2];
  double ax_by = a[0] * b[1], ay_bx = a[1] * b[0];
  o[0] = ay_bz - az_by;
  o[1] = az_bx - ax_bz;
  o[2] = ax_by - ay_bx;
}

/* triangle_intersector<double> (lbvh.h:2055-2075) on a triangle<double>
 * record {p0, e0, e1, normal} and a ray<double> {pos, dir, tmin, tmax}. */
static int
isect_reference_f64(const double* tri, const double* ray, double* t_out, double* u_out,
                    double* v_out)
{
  const double* p0 = tri;
  const double* e0 = tri + 3;
  const double* e1 = tri + 6;
  const double* nrm = tri + 9;
  const double* pos = ray;
  const double* dir = ray + 3;
  double c[3] = { p0[0] - pos[0], p0[1] - pos[1], p0[2] - pos[2] };
  double r[3];
  dcross3(dir, c, r);
  double inv_det = 1.0 / ddot3(nrm, dir);
  double u = ddot3(r, e1) * inv_det;
  double v = ddot3(r, e0) * inv_det;
  double w = (1.0 - u) - v;
  if ((u >= 0.0) && (v >= 0.0) && (w >= 0.0)) {
    double t = ddot3(nrm, c) * inv_det;
    if ((t >= ray[6]) && (t <= ray[7])) {
      *t_out = t;
      *u_out = u;
      *v_out = v;
      return 1;
    }
  }
  return 0;
}

static inline int
better64(int mode, double t, uint64_t id, double best_t, uint64_t best_id)
{
  if (mode == ORC_ISECT_REFERENCE)
    return t < best_t;
  return (t < best_t) || (t == best_t && id < best_id);
}

static void
trace_one_f64(const orc_node64* nodes, const double* tris, const double* ray, int mode,
              uint64_t* prim_out, double* t_out, double* uv_out)
{
  uint64_t st_node[ORC_STACK];
  double st_t[ORC_STACK];
  int sp = 0;
  st_node[sp] = 0;
  st_t[sp] = INFINITY;
  sp++;
  double rcp[3], inv_pos[3];
  int oct[3], inv_oct[3];
  for (int k = 0; k < 3; k++) {
    rcp[k] = 1.0 / ray[3 + k];
    double np = -ray[k];
    inv_pos[k] = np * rcp[k];
  }
  oct[0] = (ray[3] > 0.0) ? 0 : 3;
  oct[1] = (ray[4] > 0.0) ? 1 : 4;
  oct[2] = (ray[5] > 0.0) ? 2 : 5;
  inv_oct[0] = 3 - oct[0];
  inv_oct[1] = 5 - oct[1];
  inv_oct[2] = 7 - oct[2];
  uint64_t best = UINT64_MAX;
  double best_t = INFINITY, best_u = 0.0, best_v = 0.0;
  while (sp > 0) {
    sp--;
    uint64_t ni = st_node[sp];
    double etmin = st_t[sp];
    if (best_t < etmin)
      continue;
    const orc_node64* nd = &nodes[ni];
    double lt = 0.0, rt = 0.0;
    int lh = 0, rh = 0;
    for (int side = 0; side < 2; side++) {
      uint64_t ch = side ? nd->right : nd->left;
      if (ch & ORC_LEAF_BIT64) {
        uint64_t p = ch & ~ORC_LEAF_BIT64;
        double t, u, v;
        if (isect_reference_f64(tris + 12 * p, ray, &t, &u, &v) &&
            better64(mode, t, p, best_t, best)) {
          best = p;
          best_t = t;
          best_u = u;
          best_v = v;
        }
      } else {
        const orc_node64* c = &nodes[ch];
        const double b[6] = { c->min[0], c->min[1], c->min[2], c->max[0], c->max[1], c->max[2] };
        double tn[3], tf[3];
        for (int k = 0; k < 3; k++) {
          double pn = b[oct[k]] * rcp[k];
          tn[k] = pn + inv_pos[k];
          double pf = b[inv_oct[k]] * rcp[k];
          tf[k] = pf + inv_pos[k];
        }
        double tmin = dm_max(dm_max(tn[0], tn[1]), tn[2]);
        double tmax = dm_min(dm_min(tf[0], tf[1]), tf[2]);
        int h = tmax >= dm_max(0.0, tmin);
        if (side) {
          rh = h;
          rt = tmin;
        } else {
          lh = h;
          lt = tmin;
        }
      }
    }
    /* push far then near; a push onto a full stack is dropped */
#define PUSH64(N, T)                                                           \
  do {                                                                         \
    if (sp < ORC_STACK) {                                                      \
      st_node[sp] = (N);                                                       \
      st_t[sp] = (T);                                                          \
      sp++;                                                                    \
    }                                                                          \
  } while (0)
    if (lh && rh) {
      if (lt < rt) {
        PUSH64(nd->right, rt);
        PUSH64(nd->left, lt);
      } else {
        PUSH64(nd->left, lt);
        PUSH64(nd->right, rt);
      }
    } else if (lh) {
      PUSH64(nd->left, lt);
    } else if (rh) {
      PUSH64(nd->right, rt);
    }
#undef PUSH64
  }
  *prim_out = best;
  *t_out = best_t;
  if (uv_out) {
    uv_out[0] = best != UINT64_MAX ? best_u : 0.0;
    uv_out[1] = best != UINT64_MAX ? best_v : 0.0;
  }
}

typedef struct
{
  const orc_node64* nodes;
  const double* tris;
  uint64_t ntris;
  const double* rays;
  uint64_t nrays;
  int mode;
  int brute;
  uint64_t* prim;
  double* t;
  double* uv;
  int tid, nth;
} orc_job64;

static void*
job64_main(void* arg)
{
  orc_job64* j = (orc_job64*)arg;
  for (uint64_t i = (uint64_t)j->tid; i < j->nrays; i += (uint64_t)j->nth) {
    const double* ray = j->rays + 8 * i;
    if (!j->brute) {
      trace_one_f64(j->nodes, j->tris, ray, j->mode, &j->prim[i], &j->t[i],
                    j->uv ? j->uv + 2 * i : NULL);
      continue;
    }
    /* every triangle, lowest id on equal distance */
    uint64_t best = UINT64_MAX;
    double best_t = INFINITY;
    for (uint64_t p = 0; p < j->ntris; p++) {
      double t, u, v;
      if (isect_reference_f64(j->tris + 12 * p, ray, &t, &u, &v) &&
          better64(ORC_ISECT_REFERENCE_LOWID, t, p, best_t, best)) {
        best = p;
        best_t = t;
      }
    }
    j->prim[i] = best;
    j->t[i] = best_t;
  }
  return NULL;
}

static void
run_jobs64(orc_job64* proto, int nthreads)
{
  if (nthreads < 1)
    nthreads = 1;
  if (nthreads > 256)
    nthreads = 256;
  orc_job64* jobs = (orc_job64*)calloc((size_t)nthreads, sizeof(orc_job64));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int i = 0; i < nthreads; i++) {
    jobs[i] = *proto;
    jobs[i].tid = i;
    jobs[i].nth = nthreads;
  }
  for (int i = 1; i < nthreads; i++)
    pthread_create(&th[i], NULL, job64_main, &jobs[i]);
  job64_main(&jobs[0]);
  for (int i = 1; i < nthreads; i++)
    pthread_join(th[i], NULL);
  free(jobs);
  free(th);
}

void
orc_trace_f64(const orc_node64* nodes, uint64_t n, const double* tri96, const double* rays64,
              uint64_t nrays, int mode, int nthreads, uint64_t* prim, double* t, double* uv)
{
  orc_job64 j;
  memset(&j, 0, sizeof(j));
  j.nodes = nodes;
  j.tris = tri96;
  j.ntris = n;
  j.rays = rays64;
  j.nrays = nrays;
  j.mode = mode;
  j.prim = prim;
  j.t = t;
  j.uv = uv;
  run_jobs64(&j, nthreads);
}

void
orc_brute_closest_f64(const double* tri96, uint64_t n, const double* rays64, uint64_t nrays,
                      int nthreads, uint64_t* prim, double* t)
{
  orc_job64 j;
  memset(&j, 0, sizeof(j));
  j.tris = tri96;
  j.ntris = n;
  j.rays = rays64;
  j.nrays = nrays;
  j.brute = 1;
  j.prim = prim;
  j.t = t;
  run_jobs64(&j, nthreads);
}

/* FNV-1a 64-bit over raw bytes: fixture hashes of large arrays. */
uint64_t
orc_fnv1a64(const void* data, uint64_t len)
{
  const unsigned char* p = (const unsigned char*)data;
  uint64_t h = 0xcbf29ce484222325ull;
  for (uint64_t i = 0; i < len; i++) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}
