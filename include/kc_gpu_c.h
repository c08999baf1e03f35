/*
 * kc_gpu_c.h — C-ABI twins of the reference's GPU stage functions, for FFI
 * callers (ctypes, cgo, JNI). Same arguments as the C++ declarations in
 * kc_gpu.h with references turned into pointers; every function returns 0
 * on success or a non-zero status, with kc_last_error() describing it.
 *
 *   kc_initGPU                 initGPU                  src/kc_gpu.h:48, src/kc.cu:30-53
 *   kc_kmax_core_gpu           kmax_core_gpu            src/kc_gpu.h:53, src/kc.cu:770-884
 *   kc_count_deg_with_mask     count_deg_with_mask      src/kc_gpu.h:63-64, src/kc.cu:1405-1476
 *   kc_build_nbr_direct        build_nbr_direct         src/kc_gpu.h:65-66, src/kc.cu:1523-1580
 *   kc_build_graph_from_direct build_graph_from_direct  src/kc_gpu.h:67-68, src/kc.cu:1623-1673
 *   kc_tricount_direct         tricount_direct          src/kc_gpu.h:50-51, src/kc.cu:584-651
 *   kc_kmax_truss_GPU          kmax_truss_GPU           src/kc_gpu.h:55-57, src/kc.cu:1140-1309
 *   kc_test_cuda               test_cuda                src/kc_gpu.h:69
 * Offsets are 32-bit (the reference's Bint_t); the whole-graph API in
 * ktruss.h takes 64-bit offsets and device-resident inputs.
 */
#ifndef KC_AMD_KC_GPU_C_H
#define KC_AMD_KC_GPU_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kc_link {
  uint32_t x, t, d; /* vertex, position among kept vertices, degree */
} kc_link;

const char* kc_last_error(void);
int kc_initGPU(uint32_t edge_num, uint32_t N);
int kc_kmax_core_gpu(uint32_t N, const uint32_t* nbr_u, const uint32_t* nbr,
                     const uint32_t* deg_origin, uint32_t* mask, uint32_t launch_start,
                     uint32_t* result);
/* deg and deg_direct: caller-allocated, at least as many entries as kept
 * vertices (N suffices). */
int kc_count_deg_with_mask(const uint32_t* nbr, const uint32_t* nbr_u, uint32_t N,
                           const uint32_t* node_delete, uint32_t* deg_with_mask, uint32_t* to,
                           kc_link* li, uint32_t* new_N, uint32_t* deg, uint32_t* deg_direct);
int kc_build_nbr_direct(const uint32_t* nbr, const uint32_t* nbr_u, uint32_t N,
                        const uint32_t* node_delete, const uint32_t* to,
                        const uint32_t* nbr_u_direct, uint32_t* nbr_direct, uint32_t new_N);
int kc_build_graph_from_direct(const uint32_t* nbr_direct, const uint32_t* nbr_u_direct,
                               uint32_t N, uint32_t* nbr, uint32_t* fa, uint32_t* left_son,
                               uint32_t* right_son);
int kc_tricount_direct(const uint32_t* nbr_u_direct, const uint32_t* nbr_direct, uint32_t N,
                       uint32_t* nbr_tc, const uint32_t* left_son, const uint32_t* right_son,
                       uint32_t* nbr_tc_double, uint32_t tc_limit);
int kc_kmax_truss_GPU(uint32_t* k_max, uint32_t* k_max_num, uint32_t ans, uint32_t N, uint32_t k,
                      uint32_t k_step, const uint32_t* nbr_u_direct, const uint32_t* nbr_direct);
void kc_test_cuda(void);

#ifdef __cplusplus
}
#endif
#endif
