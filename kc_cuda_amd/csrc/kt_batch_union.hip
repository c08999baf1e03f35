// Batched solve as one disjoint union (no reference counterpart; the
// reference solves one graph per process, src/kc_gpu.cpp:209-347).
//
// The valid graphs of a batch are relabelled into one union graph (graph g's
// vertex v becomes vstart[g] + v), whose truss decomposition the single-graph
// solver computes in one pass (Solver::decompose_pairs_device). Truss numbers
// do not depend on the other components, so graph g's kmax is the largest
// truss number among its edges and its kmax-truss is the set of its edges
// with that truss number; a graph whose largest truss number is below 3 has
// no triangle and reports (0, 0), as the single-graph solve does.
#include "kt_batch_union.h"
#include "kt_device.h"

namespace kt {
namespace {

// graph of position x in a non-decreasing offset array of count + 1 entries:
// the last g with start[g] <= x (graphs with no entries are skipped)
template <typename T>
__device__ __forceinline__ uint32_t graph_of(const T* __restrict__ start, uint32_t count, T x) {
  uint32_t lo = 0, hi = count;  // answer in [0, count)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// status[g] = 1 when a pair of graph g is not canonical (u < v < n) or not
// strictly above its predecessor in the graph
__global__ __launch_bounds__(kBlock) void k_batch_validate(const uint32_t* __restrict__ pairs,
                                                           const uint64_t* __restrict__ poff,
                                                           const uint32_t* __restrict__ ns, uint32_t count,
                                                           uint32_t* status) {
  const uint64_t total = poff[count];
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t g = graph_of(poff, count, i);
    const uint32_t u = pairs[2 * i], v = pairs[2 * i + 1];
    bool bad = u >= v || v >= ns[g];
    if (i > poff[g]) {
      const uint32_t pu = pairs[2 * i - 2], pv = pairs[2 * i - 1];
      bad = bad || pu > u || (pu == u && pv >= v);
    }
    if (bad) status[g] = 1;
  }
}

// union pairs of the valid graphs (status 0): graph g's pairs shifted by
// vstart[g], written from position upoff[g]
__global__ __launch_bounds__(kBlock) void k_batch_union(const uint32_t* __restrict__ pairs,
                                                        const uint64_t* __restrict__ poff, uint32_t count,
                                                        const uint32_t* __restrict__ vstart,
                                                        const uint32_t* __restrict__ status,
                                                        const uint64_t* __restrict__ upoff, uint32_t* out) {
  const uint64_t total = poff[count];
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t g = graph_of(poff, count, i);
    if (status[g]) continue;
    const uint64_t j = upoff[g] + (i - poff[g]);
    out[2 * j] = pairs[2 * i] + vstart[g];
    out[2 * j + 1] = pairs[2 * i + 1] + vstart[g];
  }
}

__device__ __forceinline__ uint32_t truss_of(const uint32_t* __restrict__ del_thr, uint32_t e, uint32_t kall) {
  const uint32_t d = del_thr[e];
  return d == kNone ? kall : d + 1u;
}

// gk[g] = largest truss number among graph g's core edges
__global__ __launch_bounds__(kBlock) void k_batch_truss_max(const uint32_t* __restrict__ src0,
                                                            const uint32_t* __restrict__ inv,
                                                            const uint32_t* __restrict__ del_thr, uint32_t mc0,
                                                            const uint32_t* __restrict__ vstart, uint32_t count,
                                                            uint32_t kall, uint32_t* gk) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc0; e += gridDim.x * blockDim.x) {
    const uint32_t g = graph_of(vstart, count, inv[src0[e]]);
    const uint32_t t = truss_of(del_thr, e, kall);
    // a graph's edges all hit one counter: most see it already at least t
    // (the maximum only grows, so a stale read costs one extra atomic)
    if (t > __atomic_load_n(&gk[g], __ATOMIC_RELAXED)) atomicMax(&gk[g], t);
  }
}

// gcnt[g] = graph g's core edges with truss number gk[g] (>= 3); keep (if
// given) marks their input pairs, found by binary search in the graph's
// sorted pairs
__global__ __launch_bounds__(kBlock) void k_batch_truss_count(
    const uint32_t* __restrict__ src0, const uint32_t* __restrict__ dst0, const uint32_t* __restrict__ inv,
    const uint32_t* __restrict__ del_thr, uint32_t mc0, const uint32_t* __restrict__ vstart, uint32_t count,
    uint32_t kall, const uint32_t* __restrict__ gk, unsigned long long* gcnt, const uint32_t* __restrict__ pairs,
    const uint64_t* __restrict__ poff, uint8_t* keep) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc0; e += gridDim.x * blockDim.x) {
    const uint32_t a = inv[src0[e]], b = inv[dst0[e]];
    const uint32_t g = graph_of(vstart, count, a);
    const uint32_t k = gk[g];
    if (k < 3 || truss_of(del_thr, e, kall) != k) continue;
    atomicAdd(&gcnt[g], 1ull);
    if (!keep) continue;
    const uint32_t u = min(a, b) - vstart[g], v = max(a, b) - vstart[g];
    uint64_t lo = poff[g], hi = poff[g + 1];  // first pair >= (u, v)
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      const uint32_t pu = pairs[2 * mid], pv = pairs[2 * mid + 1];
      if (pu < u || (pu == u && pv < v))
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo < poff[g + 1] && pairs[2 * lo] == u && pairs[2 * lo + 1] == v) keep[lo] = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_batch_results(const uint32_t* __restrict__ gk,
                                                          const unsigned long long* __restrict__ gcnt,
                                                          const uint32_t* __restrict__ status, uint32_t count,
                                                          uint32_t* kmax, unsigned long long* edges,
                                                          uint32_t* status_out) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
    const bool ok = status[g] == 0 && gk[g] >= 3;
    kmax[g] = ok ? gk[g] : 0u;
    edges[g] = ok ? gcnt[g] : 0ull;
    status_out[g] = status[g];
  }
}

// (u, v, t), u < v input ids, for every core edge of truss number t >= 3
__global__ __launch_bounds__(kBlock) void k_truss_triples(const uint32_t* __restrict__ src0,
                                                          const uint32_t* __restrict__ dst0,
                                                          const uint32_t* __restrict__ inv,
                                                          const uint32_t* __restrict__ del_thr, uint32_t mc0,
                                                          uint32_t kall, uint32_t* out, uint32_t* count) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc0; e += gridDim.x * blockDim.x) {
    const uint32_t t = truss_of(del_thr, e, kall);
    if (t < 3) continue;
    const uint32_t a = inv[src0[e]], b = inv[dst0[e]];
    const uint32_t at = atomicAdd(count, 1u);
    out[3 * uint64_t(at)] = min(a, b);
    out[3 * uint64_t(at) + 1] = max(a, b);
    out[3 * uint64_t(at) + 2] = t;
  }
}

__global__ void k_fill_u8(uint8_t* p, uint64_t n, uint8_t v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

}  // namespace

void batch_validate(const uint32_t* pairs, const uint64_t* poff, const uint32_t* ns, uint32_t count,
                    uint64_t total, uint32_t* status, hipStream_t s) {
  if (total) k_batch_validate<<<grid_for(total), kBlock, 0, s>>>(pairs, poff, ns, count, status);
}

void batch_union(const uint32_t* pairs, const uint64_t* poff, uint32_t count, uint64_t total, const uint32_t* vstart,
                 const uint32_t* status, const uint64_t* upoff, uint32_t* out, hipStream_t s) {
  if (total) k_batch_union<<<grid_for(total), kBlock, 0, s>>>(pairs, poff, count, vstart, status, upoff, out);
}

void batch_truss(const CoreView& cv, uint32_t kall, const uint32_t* vstart, uint32_t count, uint32_t* gk,
                 unsigned long long* gcnt, const uint32_t* pairs, const uint64_t* poff, uint8_t* keep,
                 hipStream_t s) {
  if (!cv.mc0) return;
  k_batch_truss_max<<<grid_for(cv.mc0), kBlock, 0, s>>>(cv.src0, cv.inv, cv.del_thr, cv.mc0, vstart, count, kall,
                                                       gk);
  k_batch_truss_count<<<grid_for(cv.mc0), kBlock, 0, s>>>(cv.src0, cv.adj_d0, cv.inv, cv.del_thr, cv.mc0, vstart,
                                                         count, kall, gk, gcnt, pairs, poff, keep);
}

void batch_results(const uint32_t* gk, const unsigned long long* gcnt, const uint32_t* status, uint32_t count,
                   uint32_t* kmax, unsigned long long* edges, uint32_t* status_out, hipStream_t s) {
  if (count) k_batch_results<<<grid_for(count), kBlock, 0, s>>>(gk, gcnt, status, count, kmax, edges, status_out);
}

void truss_triples(const CoreView& cv, uint32_t kall, uint32_t* out, uint32_t* count, hipStream_t s) {
  if (cv.mc0)
    k_truss_triples<<<grid_for(cv.mc0), kBlock, 0, s>>>(cv.src0, cv.adj_d0, cv.inv, cv.del_thr, cv.mc0, kall, out,
                                                        count);
}

void fill_u8(uint8_t* p, uint64_t n, uint8_t v, hipStream_t s) {
  if (n) k_fill_u8<<<grid_for(n), kBlock, 0, s>>>(p, n, v);
}

}  // namespace kt
