// Batched k-max-truss: thousands of independent small graphs in one
// persistent launch (no reference counterpart; the reference solves one
// graph per process, src/kc_gpu.cpp:209-347).
//
// Every workgroup takes the next graph from a work counter (graphs ordered
// largest first, so the long ones start early and the short ones fill the
// tail) and solves it alone, between workgroup barriers, in a scratch slab of
// its own:
//   1. validate the pairs (canonical u < v < n, strictly sorted) and build
//      the out-rows (the pairs themselves: sorted by construction) and the
//      in-lists (pair ids grouped by their larger endpoint);
//   2. exact supports: for each edge, walk the smaller endpoint's
//      neighbourhood and look the other side up by binary search in a sorted
//      out-row (which also yields the partner edge's id);
//   3. truss levels upward from threshold 1, each peeled in
//      round-synchronous frontier rounds with the same stamps and charging
//      rule as the single-graph path (src/kc.cu:122-146), jumping to the
//      minimum live support after each level;
//   4. kmax, |T| and a keep flag per input pair.
// Lanes per edge come per graph from an autotuned table; per-graph clock
// counts feed the autotuner.
#include "kt_batch.h"
#include "kt_device.h"

namespace kt {
namespace {

constexpr int kBB = 256;  // threads per workgroup

struct Shared {
  uint32_t g, bad, qc[2], red[kBB / kWave];
};

__device__ __forceinline__ uint32_t find_edge(const uint32_t* __restrict__ P, const uint32_t* __restrict__ outs,
                                              uint32_t a, uint32_t b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  uint32_t l = outs[lo], h = outs[lo + 1];
  const uint32_t end = h;
  while (l < h) {
    const uint32_t mid = (l + h) >> 1;
    if (P[2 * mid + 1] < hi)
      l = mid + 1;
    else
      h = mid;
  }
  return (l < end && P[2 * l + 1] == hi) ? l : kNone;
}

struct GraphView {
  const uint32_t* P;  // pairs of this graph (local pair ids)
  const uint32_t *outs, *ins, *inl;
  __device__ uint32_t degree(uint32_t x) const { return (ins[x + 1] - ins[x]) + (outs[x + 1] - outs[x]); }
  // s-th neighbour of x (in-neighbours first): the other endpoint and the pair id
  __device__ void nbr(uint32_t x, uint32_t s, uint32_t* w, uint32_t* e) const {
    const uint32_t din = ins[x + 1] - ins[x];
    if (s < din) {
      const uint32_t f = inl[ins[x] + s];
      *e = f;
      *w = P[2 * f];
    } else {
      const uint32_t f = outs[x] + (s - din);
      *e = f;
      *w = P[2 * f + 1];
    }
  }
};

// exclusive scan of a[0..len) in place (a[len-1] becomes the total of the
// first len-1 entries); blockDim == kBB
__device__ void block_scan(uint32_t* a, uint32_t len, Shared& sh) {
  const uint32_t per = (len + kBB - 1) / kBB;
  const uint32_t b = threadIdx.x * per, e = min(len, b + per);
  uint32_t sum = 0;
  for (uint32_t i = b; i < e; ++i) sum += a[i];
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, &total);
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t v = a[i];
    a[i] = run;
    run += v;
  }
  __syncthreads();
}

__device__ uint32_t block_min(uint32_t v, Shared& sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, uint32_t(__shfl_xor(v, o)));
  if (lane64() == 0) sh.red[threadIdx.x / kWave] = v;
  __syncthreads();
  uint32_t r = sh.red[0];
#pragma unroll
  for (int k = 1; k < kBB / kWave; ++k) r = min(r, sh.red[k]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint32_t group_sum_rt(uint32_t v, uint32_t G) {
  for (uint32_t o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, int(o), int(G));
  return v;
}

__device__ void solve_one(const BatchArgs& A, uint32_t g, uint32_t* slab, Shared& sh) {
  const uint32_t tid = threadIdx.x;
  const uint32_t n = A.ns[g];
  const uint64_t base = A.poff[g];
  const uint32_t m = uint32_t(A.poff[g + 1] - base);
  const uint32_t* P = A.pairs + 2 * base;
  const uint32_t G = A.lanes ? A.lanes[g] : 16u;
  uint32_t* outs = slab;
  uint32_t* ins = outs + n + 1;
  uint32_t* cur = ins + n + 1;
  uint32_t* inl = cur + n;
  int32_t* sup = reinterpret_cast<int32_t*>(inl + m);
  uint32_t* st = reinterpret_cast<uint32_t*>(sup + m);
  uint32_t* del = st + m;
  uint32_t* q[2] = {del + m, del + 2 * uint64_t(m)};

  // 1. validate, count, build
  for (uint32_t i = tid; i <= n; i += kBB) outs[i] = ins[i] = 0;
  if (tid == 0) sh.bad = 0;
  __syncthreads();
  for (uint32_t e = tid; e < m; e += kBB) {
    const uint32_t u = P[2 * e], v = P[2 * e + 1];
    bool bad = u >= v || v >= n;
    if (e > 0) {
      const uint32_t pu = P[2 * e - 2], pv = P[2 * e - 1];
      bad = bad || pu > u || (pu == u && pv >= v);
    }
    if (bad) {
      atomicOr(&sh.bad, 1u);
    } else {
      atomicAdd(&outs[u], 1u);
      atomicAdd(&ins[v], 1u);
    }
  }
  __syncthreads();
  if (sh.bad) {  // uniform
    if (tid == 0) {
      A.kmax[g] = 0;
      A.count[g] = 0;
      A.status[g] = 1;
    }
    if (A.keep)
      for (uint32_t e = tid; e < m; e += kBB) A.keep[base + e] = 0;
    return;
  }
  block_scan(outs, n + 1, sh);
  block_scan(ins, n + 1, sh);
  for (uint32_t i = tid; i < n; i += kBB) cur[i] = ins[i];
  __syncthreads();
  for (uint32_t e = tid; e < m; e += kBB) inl[atomicAdd(&cur[P[2 * e + 1]], 1u)] = e;
  __syncthreads();
  const GraphView V{P, outs, ins, inl};
  const uint32_t lane = tid % G, grp = tid / G, ngrp = kBB / G;

  // 2. exact supports
  for (uint32_t e0 = 0; e0 < m; e0 += ngrp) {
    const uint32_t e = e0 + grp;
    uint32_t c = 0;
    if (e < m) {
      const uint32_t u = P[2 * e], v = P[2 * e + 1];
      const bool ux = V.degree(u) <= V.degree(v);
      const uint32_t x = ux ? u : v, y = ux ? v : u, dx = V.degree(x);
      for (uint32_t s = lane; s < dx; s += G) {
        uint32_t w, f;
        V.nbr(x, s, &w, &f);
        if (w != y && find_edge(P, outs, y, w) != kNone) ++c;
      }
    }
    c = group_sum_rt(c, G);
    if (e < m && lane == 0) {
      sup[e] = int32_t(c);
      st[e] = kNone;
      del[e] = kNone;
    }
  }
  __syncthreads();

  // 3. levels
  uint64_t alive = m;
  uint32_t r = 1, level_k = 0;
  int32_t thr = 1, floor_thr = -1;
  uint64_t level_cnt = 0;
  for (;;) {
    if (tid == 0) sh.qc[0] = sh.qc[1] = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < m; b0 += kBB) {
      const uint32_t e = b0 + tid;
      const bool p = e < m && st[e] == kNone && sup[e] < thr;
      wave_append<uint32_t>(p, e, q[0], &sh.qc[0]);
    }
    __syncthreads();
    uint32_t c = sh.qc[0];
    if (c == alive) break;  // this level would be empty: the last one stands
    int qi = 0;
    while (c) {
      for (uint32_t i = tid; i < c; i += kBB) {
        const uint32_t e = q[qi][i];
        st[e] = r;
        del[e] = uint32_t(thr);
      }
      __syncthreads();
      uint32_t* qn = q[qi ^ 1];
      uint32_t* qnc = &sh.qc[qi ^ 1];
      for (uint32_t i0 = 0; i0 < c; i0 += ngrp) {
        const uint32_t i = i0 + grp;
        const bool valid = i < c;
        uint32_t e = 0, x = 0, y = 0, dx = 0;
        if (valid) {
          e = q[qi][i];
          const uint32_t u = P[2 * e], v = P[2 * e + 1];
          const bool ux = V.degree(u) <= V.degree(v);
          x = ux ? u : v;
          y = ux ? v : u;
          dx = V.degree(x);
        }
        // one trip count per wave so the appends stay wave-aggregated
        uint32_t len = dx;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) len = max(len, uint32_t(__shfl_xor(len, o)));
        for (uint32_t s0 = 0; s0 < len; s0 += G) {
          const uint32_t s = s0 + lane;
          uint32_t xa = kNone, xb = kNone;
          if (valid && s < dx) {
            uint32_t w, exw;
            V.nbr(x, s, &w, &exw);
            const uint32_t eyw = w != y ? find_edge(P, outs, y, w) : kNone;
            if (eyw != kNone) {
              const uint32_t ma = st[exw], mb = st[eyw];
              if (ma >= r && mb >= r) {  // neither side deleted in an earlier round
                const bool ina = ma == r, inb = mb == r;
                if (!ina && !inb) {
                  xa = exw;
                  xb = eyw;
                } else if (ina && !inb) {
                  if (e < exw) xb = eyw;
                } else if (!ina && inb) {
                  if (e < eyw) xa = exw;
                }
              }
            }
          }
          const bool pa = xa != kNone && atomicSub(&sup[xa], 1) == thr;
          const bool pb = xb != kNone && atomicSub(&sup[xb], 1) == thr;
          wave_append<uint32_t>(pa, xa, qn, qnc);
          wave_append<uint32_t>(pb, xb, qn, qnc);
        }
      }
      __syncthreads();
      alive -= c;
      ++r;
      const uint32_t cn = *qnc;
      __syncthreads();
      if (tid == 0) sh.qc[qi] = 0;
      qi ^= 1;
      c = cn;
      __syncthreads();
    }
    // the level is complete: its minimum live support sets kmax and the next threshold
    uint32_t mine = kNone;
    for (uint32_t e = tid; e < m; e += kBB)
      if (st[e] == kNone) mine = min(mine, uint32_t(sup[e]));
    const uint32_t ms = block_min(mine, sh);
    if (ms == kNone) break;  // defensive: nothing live
    level_k = ms + 2;
    level_cnt = alive;
    floor_thr = thr;
    thr = int32_t(ms) + 1;
  }
  if (tid == 0) {
    A.kmax[g] = level_k;
    A.count[g] = level_k ? level_cnt : 0;
    A.status[g] = 0;
  }
  if (A.keep)
    for (uint32_t e = tid; e < m; e += kBB)
      A.keep[base + e] = (level_k && (del[e] == kNone || int64_t(del[e]) > int64_t(floor_thr))) ? 1 : 0;
}

__global__ __launch_bounds__(kBB) void k_batch_solve(BatchArgs A) {
  __shared__ Shared sh;
  uint32_t* slab = A.slab + uint64_t(blockIdx.x) * A.slab_words;
  for (;;) {
    if (threadIdx.x == 0) sh.g = atomicAdd(A.work, 1u);
    __syncthreads();
    const uint32_t idx = sh.g;
    __syncthreads();
    if (idx >= A.G) break;
    const uint32_t g = A.order[idx];
    const uint64_t t0 = wall_clock64();
    solve_one(A, g, slab, sh);
    __syncthreads();
    if (A.ticks && threadIdx.x == 0) A.ticks[g] = wall_clock64() - t0;
  }
}

__global__ void k_batch_reset(uint32_t* work) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *work = 0;
}

}  // namespace

int batch_block() { return kBB; }

uint64_t batch_slab_words(uint32_t max_n, uint64_t max_m) { return 3ull * max_n + 2 + 6ull * max_m; }

void batch_launch(const BatchArgs& a, int grid, hipStream_t s) {
  k_batch_reset<<<1, kWave, 0, s>>>(a.work);
  k_batch_solve<<<grid, kBB, 0, s>>>(a);
}

}  // namespace kt
