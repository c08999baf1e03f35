// Batched k-max-truss: thousands of independent small graphs in one
// persistent launch (no reference counterpart; the reference solves one
// graph per process, src/kc_gpu.cpp:209-347).
//
// Every workgroup takes the next graph from a work counter (graphs ordered
// largest first, so long solves start early and short ones fill the tail) and
// solves it alone, between workgroup barriers, in a scratch slab of its own,
// with the two passes of the single-graph path (src/kc_gpu.cpp:129-208,
// 336-341):
//   1. validate the pairs (canonical u < v < n, strictly sorted) and build
//      out-rows (the pairs themselves, sorted by construction) and in-lists;
//   2. core numbers by level peeling (each level jumps to the minimum alive
//      degree);
//   3. pass 1 on the max-core, pass 2 on the (L-1)-core with L = max(k^, 3)
//      and the reference's early exit (src/kc.cu:1240-1241). Each pass builds
//      a compacted adjacency of its vertex set, counts exact supports (walk
//      the smaller side's list, find the partner edge by binary search in a
//      sorted out-row) and peels truss levels in round-synchronous frontier
//      rounds with the single-graph path's stamps and charging rule
//      (src/kc.cu:122-146);
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
  unsigned long long red64[kBB / kWave];
};

// pair id of edge {a, b} (a != b) or kNone: binary search in the sorted out-row of min(a, b)
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

// exclusive scan of a[0..len) in place; blockDim == kBB
__device__ void block_scan(uint32_t* a, uint32_t len) {
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

__device__ unsigned long long block_sum64(unsigned long long v, Shared& sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane64() == 0) sh.red64[threadIdx.x / kWave] = v;
  __syncthreads();
  unsigned long long r = 0;
#pragma unroll
  for (int k = 0; k < kBB / kWave; ++k) r += sh.red64[k];
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint32_t group_sum_rt(uint32_t v, uint32_t G) {
  for (uint32_t o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, int(o), int(G));
  return v;
}

// Scratch of one graph, carved from the workgroup's slab.
struct Work {
  const uint32_t* P;
  uint32_t n, m, G;
  uint32_t *outs, *ins, *cur, *core, *vq[2], *soff, *inl;
  int32_t* sup;
  uint32_t *st, *del, *q[2], *snb, *se;
};

// s-th neighbour of x in the full graph (in-neighbours first) and its pair id
__device__ __forceinline__ void full_nbr(const Work& W, uint32_t x, uint32_t s, uint32_t* w, uint32_t* e) {
  const uint32_t din = W.ins[x + 1] - W.ins[x];
  if (s < din) {
    const uint32_t f = W.inl[W.ins[x] + s];
    *e = f;
    *w = W.P[2 * f];
  } else {
    const uint32_t f = W.outs[x] + (s - din);
    *e = f;
    *w = W.P[2 * f + 1];
  }
}

__device__ __forceinline__ uint32_t full_deg(const Work& W, uint32_t x) {
  return (W.ins[x + 1] - W.ins[x]) + (W.outs[x + 1] - W.outs[x]);
}

// Core numbers of every vertex (core[v]); returns the largest.
__device__ uint32_t core_numbers(const Work& W, Shared& sh) {
  const uint32_t tid = threadIdx.x, n = W.n;
  uint32_t* deg = W.cur;
  for (uint32_t v = tid; v < n; v += kBB) {
    deg[v] = full_deg(W, v);
    W.core[v] = kNone;
  }
  __syncthreads();
  uint32_t alive = n, k = 1, kmax_core = 0;
  while (alive) {
    if (tid == 0) sh.qc[0] = sh.qc[1] = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < n; b0 += kBB) {
      const uint32_t v = b0 + tid;
      const bool p = v < n && W.core[v] == kNone && deg[v] < k;
      wave_append<uint32_t>(p, v, W.vq[0], &sh.qc[0]);
    }
    __syncthreads();
    uint32_t c = sh.qc[0];
    int qi = 0;
    while (c) {
      for (uint32_t i = tid; i < c; i += kBB) W.core[W.vq[qi][i]] = k - 1;
      __syncthreads();
      kmax_core = k - 1;
      uint32_t* qn = W.vq[qi ^ 1];
      uint32_t* qnc = &sh.qc[qi ^ 1];
      // one wave per removed vertex; alive neighbours lose a degree
      for (uint32_t i0 = 0; i0 < c; i0 += kBB / kWave) {
        const uint32_t i = i0 + tid / kWave;
        const bool valid = i < c;
        const uint32_t v = valid ? W.vq[qi][i] : 0;
        uint32_t d = valid ? full_deg(W, v) : 0;
        uint32_t len = d;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) len = max(len, uint32_t(__shfl_xor(len, o)));
        for (uint32_t s0 = 0; s0 < len; s0 += kWave) {
          const uint32_t s = s0 + lane64();
          bool cross = false;
          uint32_t w = 0, e;
          if (valid && s < d) {
            full_nbr(W, v, s, &w, &e);
            if (W.core[w] == kNone) cross = atomicSub(&deg[w], 1u) == k;
          }
          wave_append<uint32_t>(cross, w, qn, qnc);
        }
      }
      __syncthreads();
      alive -= c;
      const uint32_t cn = *qnc;
      __syncthreads();
      if (tid == 0) sh.qc[qi] = 0;
      qi ^= 1;
      c = cn;
      __syncthreads();
    }
    if (!alive) break;
    uint32_t mine = kNone;
    for (uint32_t v = tid; v < n; v += kBB)
      if (W.core[v] == kNone) mine = min(mine, deg[v]);
    k = block_min(mine, sh) + 1;
  }
  return kmax_core;
}

// Compacted adjacency of the vertices with core >= kc: soff / snb (neighbour) /
// se (pair id); edges inside get st = kNone, the others st = 0 (treated as
// deleted before round 1). Returns the number of edges inside.
__device__ uint64_t build_sub(const Work& W, uint32_t kc, Shared& sh) {
  const uint32_t tid = threadIdx.x, n = W.n, m = W.m;
  for (uint32_t v = tid; v <= n; v += kBB) W.soff[v] = 0;
  __syncthreads();
  unsigned long long inside = 0;
  for (uint32_t e = tid; e < m; e += kBB) {
    const uint32_t u = W.P[2 * e], v = W.P[2 * e + 1];
    const bool in = W.core[u] >= kc && W.core[v] >= kc;
    W.st[e] = in ? kNone : 0u;
    W.del[e] = kNone;
    if (in) {
      atomicAdd(&W.soff[u], 1u);
      atomicAdd(&W.soff[v], 1u);
      ++inside;
    }
  }
  __syncthreads();
  block_scan(W.soff, n + 1);
  for (uint32_t v = tid; v < n; v += kBB) W.cur[v] = W.soff[v];
  __syncthreads();
  for (uint32_t e = tid; e < m; e += kBB) {
    if (W.st[e] != kNone) continue;
    const uint32_t u = W.P[2 * e], v = W.P[2 * e + 1];
    uint32_t p = atomicAdd(&W.cur[u], 1u);
    W.snb[p] = v;
    W.se[p] = e;
    p = atomicAdd(&W.cur[v], 1u);
    W.snb[p] = u;
    W.se[p] = e;
  }
  __syncthreads();
  return block_sum64(inside, sh);
}

__device__ __forceinline__ uint32_t sub_deg(const Work& W, uint32_t x) { return W.soff[x + 1] - W.soff[x]; }

// Exact supports of the live edges (triangles with all three edges live).
__device__ void count_sub(const Work& W) {
  const uint32_t tid = threadIdx.x, m = W.m, G = W.G;
  const uint32_t lane = tid % G, grp = tid / G, ngrp = kBB / G;
  for (uint32_t e0 = 0; e0 < m; e0 += ngrp) {
    const uint32_t e = e0 + grp;
    uint32_t c = 0;
    const bool live = e < m && W.st[e] == kNone;
    if (live) {
      const uint32_t u = W.P[2 * e], v = W.P[2 * e + 1];
      const bool ux = sub_deg(W, u) <= sub_deg(W, v);
      const uint32_t x = ux ? u : v, y = ux ? v : u, dx = sub_deg(W, x), b = W.soff[x];
      for (uint32_t s = lane; s < dx; s += G) {
        const uint32_t w = W.snb[b + s];
        if (w == y) continue;
        const uint32_t f = find_edge(W.P, W.outs, y, w);
        if (f != kNone && W.st[f] == kNone) ++c;
      }
    }
    c = group_sum_rt(c, G);
    if (live && lane == 0) W.sup[e] = int32_t(c);
  }
  __syncthreads();
}

struct Level {
  uint32_t kmax;
  uint64_t count;
  int32_t floor_thr;
  uint32_t levels, rounds;
};

// Truss levels on the live edges upward from thr (a level keeps the edges of
// support >= thr); early exit on a first level of early_ans edges at kmax
// early_k. Returns the last non-empty level.
__device__ Level levels(const Work& W, uint64_t alive, int32_t thr, uint64_t early_ans, uint32_t early_k,
                        Shared& sh) {
  const uint32_t tid = threadIdx.x, m = W.m, G = W.G;
  const uint32_t lane = tid % G, grp = tid / G, ngrp = kBB / G;
  Level out{0, 0, -1, 0, 0};
  uint32_t r = 1;
  bool first = true;
  for (;;) {
    if (tid == 0) sh.qc[0] = sh.qc[1] = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < m; b0 += kBB) {
      const uint32_t e = b0 + tid;
      const bool p = e < m && W.st[e] == kNone && W.sup[e] < thr;
      wave_append<uint32_t>(p, e, W.q[0], &sh.qc[0]);
    }
    __syncthreads();
    uint32_t c = sh.qc[0];
    if (c == alive) break;  // this level would be empty: the last one stands
    int qi = 0;
    while (c) {
      for (uint32_t i = tid; i < c; i += kBB) {
        const uint32_t e = W.q[qi][i];
        W.st[e] = r;
        W.del[e] = uint32_t(thr);
      }
      __syncthreads();
      uint32_t* qn = W.q[qi ^ 1];
      uint32_t* qnc = &sh.qc[qi ^ 1];
      for (uint32_t i0 = 0; i0 < c; i0 += ngrp) {
        const uint32_t i = i0 + grp;
        const bool valid = i < c;
        uint32_t e = 0, x = 0, y = 0, dx = 0, b = 0;
        if (valid) {
          e = W.q[qi][i];
          const uint32_t u = W.P[2 * e], v = W.P[2 * e + 1];
          const bool ux = sub_deg(W, u) <= sub_deg(W, v);
          x = ux ? u : v;
          y = ux ? v : u;
          dx = sub_deg(W, x);
          b = W.soff[x];
        }
        // one trip count per wave so the appends stay wave-aggregated
        uint32_t len = dx;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) len = max(len, uint32_t(__shfl_xor(len, o)));
        for (uint32_t s0 = 0; s0 < len; s0 += G) {
          const uint32_t s = s0 + lane;
          uint32_t xa = kNone, xb = kNone;
          if (valid && s < dx) {
            const uint32_t w = W.snb[b + s], exw = W.se[b + s];
            const uint32_t eyw = w != y ? find_edge(W.P, W.outs, y, w) : kNone;
            if (eyw != kNone) {
              const uint32_t ma = W.st[exw], mb = W.st[eyw];
              if (ma >= r && mb >= r) {  // neither side deleted in an earlier round (or outside)
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
          const bool pa = xa != kNone && atomicSub(&W.sup[xa], 1) == thr;
          const bool pb = xb != kNone && atomicSub(&W.sup[xb], 1) == thr;
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
    out.levels += 1;
    out.rounds = r - 1;
    uint32_t mine = kNone;
    for (uint32_t e = tid; e < m; e += kBB)
      if (W.st[e] == kNone) mine = min(mine, uint32_t(W.sup[e]));
    const uint32_t ms = block_min(mine, sh);
    if (ms == kNone) break;  // the cascade emptied the level
    out.kmax = ms + 2;
    out.count = alive;
    out.floor_thr = thr;
    if (first && early_ans && alive == early_ans && ms + 2 == early_k) break;
    first = false;
    thr = int32_t(ms) + 1;
  }
  return out;
}

__device__ void solve_one(const BatchArgs& A, uint32_t g, uint32_t* slab, Shared& sh, uint64_t t0) {
  const uint32_t tid = threadIdx.x;
  unsigned long long* ph = A.phases ? A.phases + uint64_t(g) * kBatchPhases : nullptr;
  auto mark = [&](int i) {  // after a barrier: every thread agrees the phase is over
    if (ph && tid == 0) ph[i] = wall_clock64() - t0;
  };
  const uint32_t n = A.ns[g];
  const uint64_t base = A.poff[g];
  const uint32_t m = uint32_t(A.poff[g + 1] - base);
  Work W;
  W.P = A.pairs + 2 * base;
  W.n = n;
  W.m = m;
  W.G = A.lanes ? A.lanes[g] : 16u;
  W.outs = slab;
  W.ins = W.outs + n + 1;
  W.soff = W.ins + n + 1;
  W.cur = W.soff + n + 1;
  W.core = W.cur + n;
  W.vq[0] = W.core + n;
  W.vq[1] = W.vq[0] + n;
  W.inl = W.vq[1] + n;
  W.sup = reinterpret_cast<int32_t*>(W.inl + m);
  W.st = reinterpret_cast<uint32_t*>(W.sup + m);
  W.del = W.st + m;
  W.q[0] = W.del + m;
  W.q[1] = W.q[0] + m;
  W.snb = W.q[1] + m;
  W.se = W.snb + 2 * uint64_t(m);

  // 1. validate, count, build
  for (uint32_t i = tid; i <= n; i += kBB) W.outs[i] = W.ins[i] = 0;
  if (tid == 0) sh.bad = 0;
  __syncthreads();
  for (uint32_t e = tid; e < m; e += kBB) {
    const uint32_t u = W.P[2 * e], v = W.P[2 * e + 1];
    bool bad = u >= v || v >= n;
    if (e > 0) {
      const uint32_t pu = W.P[2 * e - 2], pv = W.P[2 * e - 1];
      bad = bad || pu > u || (pu == u && pv >= v);
    }
    if (bad) {
      atomicOr(&sh.bad, 1u);
    } else {
      atomicAdd(&W.outs[u], 1u);
      atomicAdd(&W.ins[v], 1u);
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
  block_scan(W.outs, n + 1);
  block_scan(W.ins, n + 1);
  for (uint32_t i = tid; i < n; i += kBB) W.cur[i] = W.ins[i];
  __syncthreads();
  for (uint32_t e = tid; e < m; e += kBB) W.inl[atomicAdd(&W.cur[W.P[2 * e + 1]], 1u)] = e;
  __syncthreads();
  mark(0);

  // 2. core numbers
  const uint32_t cmax = core_numbers(W, sh);
  mark(1);

  // 3. pass 1 on the max-core, pass 2 on the (L-1)-core
  Level p1{0, 0, -1, 0, 0};
  if (cmax >= 2) {
    const uint64_t m1 = build_sub(W, cmax, sh);
    count_sub(W);
    mark(2);
    p1 = levels(W, m1, 1, 0, 0, sh);
  }
  mark(3);
  const uint32_t L = max(p1.kmax, 3u);
  Level p2{0, 0, -1, 0, 0};
  if (cmax >= L - 1) {
    const uint64_t m2 = build_sub(W, L - 1, sh);
    count_sub(W);
    mark(4);
    p2 = levels(W, m2, int32_t(L) - 2, p1.kmax >= 3 ? p1.count : 0, p1.kmax, sh);
  } else {
    for (uint32_t e = tid; e < m; e += kBB) W.st[e] = 0;  // nothing survives
    __syncthreads();
  }
  mark(5);
  if (ph && tid == 0) {
    ph[6] = (uint64_t(p1.levels) << 32) | p1.rounds;
    ph[7] = (uint64_t(p2.levels) << 32) | p2.rounds;
  }

  // 4. results
  if (tid == 0) {
    A.kmax[g] = p2.kmax;
    A.count[g] = p2.kmax ? p2.count : 0;
    A.status[g] = 0;
  }
  if (A.keep)
    for (uint32_t e = tid; e < m; e += kBB) {
      const uint32_t d = W.del[e];
      A.keep[base + e] = (p2.kmax && W.st[e] != 0 && (d == kNone || int64_t(d) > int64_t(p2.floor_thr))) ? 1 : 0;
    }
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
    solve_one(A, g, slab, sh, t0);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t t1 = wall_clock64();
      if (A.ticks) A.ticks[g] = t1 - t0;
      if (A.phases) {
        A.phases[uint64_t(g) * kBatchPhases + 8] = t0;
        A.phases[uint64_t(g) * kBatchPhases + 9] = t1;
      }
    }
  }
}

__global__ void k_batch_reset(uint32_t* work) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *work = 0;
}

}  // namespace

int batch_block() { return kBB; }

uint64_t batch_slab_words(uint32_t max_n, uint64_t max_m) { return 7ull * max_n + 3 + 10ull * max_m; }

void batch_launch(const BatchArgs& a, int grid, hipStream_t s) {
  k_batch_reset<<<1, kWave, 0, s>>>(a.work);
  k_batch_solve<<<grid, kBB, 0, s>>>(a);
}

}  // namespace kt
