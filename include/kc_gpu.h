/*
 * kc_gpu.h — link-compatible declarations of the reference's GPU stage
 * functions (src/kc_gpu.h:45-69), implemented by libkc_amd.so on MI355X.
 *
 * A host program written against the reference (its src/kc_gpu.cpp) links
 * against libkc_amd.so instead of the reference's kc.cu and runs unchanged:
 * same names, C++ linkage, argument types and meanings; errors print and exit
 * with the runtime's code, as in the reference (src/kc.cu:14-23).
 *
 * Types and constants the reference's host code relies on keep their layout:
 *   Bint_t   32-bit offsets and edge ids          (src/kc_gpu.h:6)
 *   Edge_t   (u, v), 8 bytes                       (src/kc_gpu.h:26-31)
 *   Nbr_t    (v, fa), 8-byte aligned               (src/kc_gpu.h:33-38)
 *   Link_t   (x, t, d) sort record                 (src/kc_gpu.h:40-43)
 *   K_MIDDLE step of the k search                  (src/kc_gpu.h:21)
 * cmp_nbr / cmp_link are defined by the caller, as in the reference.
 */
#ifndef KC_AMD_KC_GPU_H
#define KC_AMD_KC_GPU_H

#include <cstdint>

typedef uint32_t Bint_t;

#define K_MIDDLE 19
#define K_CORE_STEP_LARGE 500
#define K_CORE_STEP_MIDDLE 25

struct Edge_t {
  uint32_t u;
  uint32_t v;
  Edge_t() {}
  Edge_t(const uint32_t& a, const uint32_t& b) : u(a), v(b) {}
} __attribute__((aligned(4)));

struct Nbr_t {
  uint32_t v;
  Bint_t fa;
  Nbr_t() {}
  Nbr_t(uint32_t w, Bint_t f) : v(w), fa(f) {}
} __attribute__((aligned(8)));

struct Link_t {
  uint32_t x, t, d;  // vertex, position among kept vertices, degree
} __attribute__((aligned(4)));

bool cmp_nbr(Nbr_t& a, Nbr_t& b);
bool cmp_link(Link_t& a, Link_t& b);

/* Device warm-up; prints memory figures when KC_VERBOSE=1. */
void initGPU(const Bint_t edge_num, const uint32_t N);

/* Supports of the oriented edges, doubled (nbr_tc[e] = 2 * sup(e)); with
 * tc_limit == 0 also copied to both slots of every edge in nbr_tc_double. */
void tricount_direct(const Bint_t* nbr_u_direct, const uint32_t* nbr_direct, const uint32_t N,
                     const Bint_t* nbr_u, const uint32_t* nbr, uint32_t* nbr_tc,
                     const Bint_t* left_son, const Bint_t* right_son, uint32_t* nbr_tc_double,
                     uint32_t tc_limit = 0);

/* Launch_start == 0xFFFFFFFF: marks (mask = 1) every vertex outside the
 * max-core and returns its k - 1; otherwise marks the vertices outside the
 * Launch_start-core and returns Launch_start - 1 (0 for Launch_start 0). */
uint32_t kmax_core_gpu(const uint32_t N, const Bint_t* nbr_u, const uint32_t* nbr,
                       const uint32_t* deg_origin, uint32_t*& mask, uint32_t Launch_start);

/* Largest truss level reachable from the doubled threshold k with first
 * level k + 2 k_step (early exit when that level keeps `ans` edges):
 * k_max = level / 2 + 2, k_max_num = its edge count; unchanged if none. */
void kmax_truss_GPU(uint32_t& k_max, Bint_t& k_max_num, Bint_t ans, uint32_t N, uint32_t k,
                    uint32_t k_step, Bint_t*& nbr_u_direct, uint32_t*& nbr_direct, Bint_t*& nbr_u,
                    uint32_t*& nbr, Bint_t*& fa, Bint_t*& left_son, Bint_t*& right_son,
                    uint32_t*& nbr_mask, uint32_t*& nbr_mask_double, uint32_t*& nbr_tc,
                    uint32_t*& nbr_tc_double, uint32_t& accurate_tc);

/* Degrees inside the kept vertices, their rank by (degree ascending, id
 * descending) and oriented out-degrees; allocates deg and deg_direct. */
void count_deg_with_mask(const uint32_t*& nbr, const Bint_t*& nbr_u, uint32_t N,
                         uint32_t*& node_delete, uint32_t*& deg_with_mask, uint32_t*& to,
                         Link_t*& li, uint32_t& new_N, uint32_t*& deg, uint32_t*& deg_direct);

/* Oriented CSR of the kept vertices in rank ids, rows ascending. */
void build_nbr_direct(const uint32_t*& nbr, const Bint_t*& nbr_u, uint32_t N,
                      uint32_t*& node_delete, uint32_t*& to, uint32_t*& deg,
                      Bint_t*& nbr_u_direct, uint32_t*& nbr_direct, uint32_t new_N);

/* Undirected CSR of an oriented CSR with the edge id of every slot (fa) and
 * each edge's slot in its lower / higher endpoint's row. */
void build_graph_from_direct(uint32_t*& nbr_direct, Bint_t*& nbr_u_direct, uint32_t N,
                             uint32_t*& nbr, Bint_t*& nbr_u, Bint_t*& fa, Bint_t*& left_son,
                             Bint_t*& right_son);

void test_cuda();

#endif
