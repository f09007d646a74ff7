/*
 * ORACLE — test infrastructure only.
 *
 * A clean-room CPU restatement of the tay10r/lbvh build and closest-hit path
 * (lbvh.h), used as the checker for the HIP engine and as the "port" CPU
 * baseline.  It is never linked into, loaded by or called from the product
 * library: only tests, the smoke check and the benchmark's CPU-baseline leg
 * use it.
 *
 * Every function states the reference lines it follows.  Floating-point
 * semantics are reproduced exactly; the file must be compiled with
 * -ffp-contract=off and without fast-math.
 *
 * Pinning (what checks each function beyond this restatement):
 *   orc_triangle_aabb      golden fixtures from the compiled lbvh.h; closed form
 *   orc_centroid_bounds    golden fixtures; numpy min/max of centres
 *   orc_morton_codes       golden fixtures; per-bit interleave over all lane values
 *   orc_sort_pairs         numpy stable argsort (library routine)
 *   orc_karras             golden fixtures; brute-force radix tree by definition;
 *                          structural invariants of lbvh_test.cpp:369-446
 *   orc_fit                golden fixtures; subtree min/max closed form;
 *                          volume invariant of lbvh_test.cpp:458-518
 *   orc_trace (modes 0,1)  golden fixtures; brute-force closest hit
 *   orc_*_f64 (builder<double>)
 *                          golden fixtures from the compiled lbvh.h
 *                          (builder<double> over triangle<double> and over
 *                          plain boxes); numpy min/max of centres; brute-force
 *                          radix tree; subtree min/max closed form
 *   orc_trace_f64          the compiled lbvh.h traverser<double> (oracle/_ref);
 *                          brute-force FP64 closest hit
 *   orc_trace (mode 2)     brute-force closest hit with the same test; FP64
 *                          plane-distance error bound.  No reference counterpart
 *                          exists for the watertight test: its hit set is pinned
 *                          against brute force and FP64 only.
 */
#ifndef LBVH_ORACLE_H
#define LBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lbvh::aabb<float> (lbvh.h:205-212): 24 bytes. */
typedef struct
{
  float min[3];
  float max[3];
} orc_aabb;

/* lbvh::node<float> (lbvh.h:220-255): 32 bytes, leaf flag in bit 31. */
typedef struct
{
  orc_aabb box;
  uint32_t left;
  uint32_t right;
} orc_node;

#define ORC_LEAF_BIT 0x80000000u

enum
{
  ORC_LEAF_REFERENCE = 0, /* leaf index = sorted position (lbvh.h:1728-1729) */
  ORC_LEAF_PRIMITIVE_ID = 1 /* leaf index = primitive id of that position */
};

enum
{
  ORC_ISECT_REFERENCE = 0,       /* triangle_intersector, first hit wins ties */
  ORC_ISECT_REFERENCE_LOWID = 1, /* same test, ties to the lowest primitive id */
  ORC_ISECT_WATERTIGHT = 2       /* watertight test, ties to the lowest id */
};

void orc_triangle_aabb(const float* tri, orc_aabb* out);
void orc_centroid_bounds(const float* tris, uint64_t n, orc_aabb* out);
void orc_morton_codes(const float* tris, uint64_t n, int code_bits,
                      const orc_aabb* bounds, void* codes);
void orc_sort_pairs(int code_bits, const void* codes, uint64_t n,
                    void* sorted_codes, uint32_t* ids);
int orc_karras(int code_bits, const void* sorted_codes, uint64_t n,
               orc_node* nodes);
void orc_fit(orc_node* nodes, uint64_t n, const float* tris,
             const uint32_t* leaf_prim);
int orc_build(const float* tris, uint64_t n, int code_bits, int leaf_order,
              orc_node* nodes, orc_aabb* bounds_out, void* sorted_codes_out,
              uint32_t* ids_out);

/* lbvh::node<double> (lbvh.h:220-255 with scalar_type = double): 64 bytes,
 * leaf flag in bit 63. */
typedef struct
{
  double min[3];
  double max[3];
  uint64_t left;
  uint64_t right;
} orc_node64;

#define ORC_LEAF_BIT64 0x8000000000000000ull

/* builder<double> over converter boxes (6 doubles each). */
void orc_triangle_aabb_f64(const double* tri96, uint64_t n, double* boxes48);
void orc_morton_codes_f64(const double* boxes48, uint64_t n, double* bounds6,
                          uint64_t* codes);
int orc_karras_f64(const uint64_t* sorted_codes, uint64_t n,
                   orc_node64* nodes);
void orc_fit_f64(orc_node64* nodes, uint64_t n, const double* boxes48);
int orc_build_f64(const double* boxes48, uint64_t n, int leaf_order,
                  orc_node64* nodes, uint64_t* sorted_codes_out,
                  uint32_t* ids_out);
/* traverser<double> with triangle_intersector<double> over triangle<double>
 * records (12 doubles) and ray<double> records (8 doubles); modes
 * ORC_ISECT_REFERENCE and ORC_ISECT_REFERENCE_LOWID; prim UINT64_MAX on a
 * miss, uv 0. */
void orc_trace_f64(const orc_node64* nodes, uint64_t n, const double* tri96,
                   const double* rays64, uint64_t nrays, int mode, int nthreads,
                   uint64_t* prim, double* t, double* uv);
void orc_brute_closest_f64(const double* tri96, uint64_t n, const double* rays64,
                           uint64_t nrays, int nthreads, uint64_t* prim, double* t);
void orc_trace(const orc_node* nodes, uint64_t n, const float* tris,
               const float* rays, uint64_t nrays, int mode, int nthreads,
               uint32_t* prim, float* t, float* uv, uint64_t* counters);
void orc_brute_closest(const float* tris, uint64_t n, const float* rays,
                       uint64_t nrays, int mode, int nthreads, uint32_t* prim,
                       float* t);
uint64_t orc_fnv1a64(const void* data, uint64_t len);
void orc_t_fp64(const float* tris, const float* rays, const uint32_t* prim,
                uint64_t nrays, double* t64);

#ifdef __cplusplus
}
#endif

#endif
