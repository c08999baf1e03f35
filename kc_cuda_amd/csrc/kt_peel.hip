// Device-driven edge-peel rounds (truss levels) on CDNA4.
//
// A level removes, round by round, every edge whose support is below the
// threshold (reference: src/kc.cu:886-1056, __delete_edge / ktruss), and the
// search moves from level to level (src/kc.cu:1140-1309). Here both run on
// the device: a round is two launches that read what to do, the frontier,
// its size, the round stamp and the threshold from the control block.
//   peel     walk the triangles of the frontier, decrement the partners,
//            append edges that cross below the threshold to the next queue;
//            at a level change: collect the new level's frontier;
//   next     clear the walked frontier from the live bit matrix (dense
//            cores); every workgroup takes the same decision on the next
//            frontier from the control block and stamps it (state, deletion
//            threshold, frontier bits), or, when it is empty, reduces the
//            live supports; the last workgroup switches queues, records a
//            finished level and sets what the next round does.
// The host enqueues rounds in batches, reads the control block once per batch
// and steps in only to drop a frontier, compact, or collect the answer; a
// kernel whose action does not concern it does nothing. k_round_stamp stamps
// the first frontier of a search (or ends the level, when it is empty).
#include <algorithm>
#include <cstdlib>

#include "kt_device.h"
#include "kt_kernels.h"

namespace kt {
namespace {

// Grids of the round kernels. Stamp and next end with one ticket atomic per
// workgroup, so they stay small; their loops stride over the frontier (or
// the edges, for a level's minimum support).
constexpr int kStampGrid = 256;
constexpr int kPeelGrid = 1024;
constexpr uint32_t kPeelUnroll = 8;      // triangles per lane per step of the dense peel
constexpr uint32_t kSparseStage = 4096;  // LDS words per workgroup for the sparse peel's staged rows
constexpr int kClearGrid = 512;

// What the next round does with a frontier of c edges when `alive` edges are
// live and unstamped.
__device__ __forceinline__ uint32_t decide_round(uint32_t c, unsigned long long alive, int32_t thr,
                                                 const RoundRule& rule, bool allow_compact) {
  if (c == 0) return kRoundMinSup;  // the level is complete
  // A non-empty level at thr is a (thr + 2)-truss: each of its vertices has
  // at least thr + 1 neighbours on it, so it has at least
  // (thr + 2)(thr + 1) / 2 edges. With fewer live edges left after this
  // frontier (none: the frontier is every live edge) the level empties:
  // every live edge dies at thr now.
  if (c == alive || (thr > 0 && alive - c < (unsigned long long)(thr + 2) * (unsigned long long)(thr + 1) / 2))
    return kRoundEmpty;
  if (allow_compact && 2ull * alive <= rule.mc) return kRoundCompact;
  const bool drop = rule.dense ? double(c) * (double(rule.ws) + rule.tri_cost * double(thr)) >=
                                     double(alive) * double(rule.ws) + rule.drop_fixed
                               : double(c) * rule.drop_ratio >= double(alive);
  return drop ? kRoundDrop : kRoundPeel;
}

__device__ __forceinline__ void set_action(Ctl* ctl, uint32_t a) {
  if (a == kRoundMinSup) ctl->minsup = kNone;
  ctl->action = a;
}

__global__ void k_round_start(Ctl* ctl, uint32_t r, unsigned long long alive, int32_t thr, RoundRule rule,
                              uint32_t reset, SearchRule search) {
  if (threadIdx.x != 0) return;
  const uint32_t c = ctl->qcount[0];
  ctl->cur = 0;
  ctl->qcount[1] = 0;
  ctl->round = r;
  ctl->alive = alive;
  ctl->thr = thr;
  ctl->ticket = 0;
  ctl->stamped = 0;
  ctl->minsup = kNone;
  if (reset) {
    ctl->err &= ~kErrPeelCount;
    ctl->rounds_done = 0;
    ctl->levels = 0;
    ctl->first = 1;
    ctl->level_kmax = 0;
    ctl->level_count = 0;
    ctl->level_floor = -1;
    ctl->early_ans = search.early_ans;
    ctl->early_k = search.early_k;
    ctl->exact_early = search.exact_early;
    ctl->max_levels = search.max_levels;
  }
  set_action(ctl, decide_round(c, alive, thr, rule, false));
}

// A level is complete (its frontier is empty): every workgroup reduces the
// supports of the live edges; the last one (after switch_queue(): the walked
// round's queue switch) records the level (kmax = min support + 2, its edge
// count and threshold), applies the early exit and the level budget, and
// moves to threshold min + 1: a sweep next, or a compaction of a sparse core
// that has lost half its edges. ctl->minsup is kNone on entry and on exit.
template <class Switch>
__device__ __forceinline__ void level_end(Ctl* ctl, const uint32_t* __restrict__ state,
                                          const int32_t* __restrict__ sup, const RoundRule& rule,
                                          Switch switch_queue) {
  __shared__ uint32_t s_min[kBlock / kWave];
  uint32_t best = kNone;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < rule.mc; e += gridDim.x * blockDim.x)
    if (state[e] == kNone) best = min(best, uint32_t(sup[e]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = min(best, uint32_t(__shfl_xor(best, o)));
  if (lane64() == 0) s_min[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / kWave; ++k) best = min(best, s_min[k]);
    if (best != kNone) {
      const uint32_t prev = atomicMin(&ctl->minsup, best);
      asm volatile("" ::"v"(prev));  // performed before this workgroup's ticket
    }
  }
  if (last_block(&ctl->ticket) && threadIdx.x == 0) {
    switch_queue();
    const uint32_t ms = ld_agent(&ctl->minsup);
    atomicExch(&ctl->minsup, kNone);
    const unsigned long long alive = ctl->alive;
    if (ms == kNone || alive == 0) {  // nothing live: the search is over
      if ((ms == kNone) != (alive == 0)) ctl->err |= kErrPeelCount;
      ctl->action = kRoundDone;
      return;
    }
    ctl->level_kmax = ms + 2;
    ctl->level_count = alive;
    ctl->level_floor = ctl->thr;
    ctl->levels += 1;
    if (ctl->first && ctl->early_ans && alive == ctl->early_ans && (ctl->exact_early || ms + 2 == ctl->early_k)) {
      ctl->level_kmax = ctl->early_k;
      ctl->action = kRoundDone;
      return;
    }
    ctl->first = 0;
    if (ctl->max_levels && ctl->levels >= ctl->max_levels) {  // level budget spent
      ctl->action = kRoundDone;
      return;
    }
    ctl->thr = int32_t(ms) + 1;
    ctl->action = (!rule.dense && 2ull * alive <= rule.mc) ? kRoundCompact : kRoundSweep;
  }
}

// Stamp of a frontier edge: state = round, deletion threshold of its base
// edge = thr, and (frontiers walked on dense cores) its frontier bits.
__device__ __forceinline__ void stamp_edge(uint32_t e, uint32_t r, uint32_t thr, uint32_t* state,
                                           const uint32_t* __restrict__ orig, uint32_t* del_thr,
                                           const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                           unsigned long long* fbits, uint32_t ws) {
  state[e] = r;
  del_thr[orig[e]] = thr;
  if (fbits) {
    const uint32_t u = src[e], v = dst[e];
    atomicOr(&fbits[uint64_t(u) * ws + (v >> 6)], 1ull << (v & 63));
    atomicOr(&fbits[uint64_t(v) * ws + (u >> 6)], 1ull << (u & 63));
  }
}

// An emptied level (kRoundEmpty): every live edge of the edge set, the
// frontier included, dies in round r at thr.
__device__ __forceinline__ void stamp_all_live(uint32_t mc, uint32_t r, uint32_t thr, uint32_t* state,
                                               const uint32_t* __restrict__ orig, uint32_t* del_thr) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc; e += gridDim.x * blockDim.x)
    if (state[e] == kNone) {
      state[e] = r;
      del_thr[orig[e]] = thr;
    }
}

// Stamp the first frontier of a search once (the rounds queued after a drop
// or an empty level find it stamped already); with an empty frontier, the
// level end instead.
__global__ __launch_bounds__(kBlock) void k_round_stamp(uint32_t* q0, uint32_t* q1, Ctl* ctl, uint32_t* state,
                                                        const int32_t* __restrict__ sup,
                                                        const uint32_t* __restrict__ orig, uint32_t* del_thr,
                                                        const uint32_t* __restrict__ src,
                                                        const uint32_t* __restrict__ dst,
                                                        unsigned long long* fbits, uint32_t ws, RoundRule rule) {
  const uint32_t a = ctl->action;
  if (a == kRoundMinSup) {
    level_end(ctl, state, sup, rule, [] {});
    return;
  }
  if (ctl->stamped || (a != kRoundPeel && a != kRoundDrop && a != kRoundEmpty)) return;
  const uint32_t cur = ctl->cur, cnt = ctl->qcount[cur], r = ctl->round;
  const uint32_t thr = uint32_t(ctl->thr);
  const uint32_t* q = cur ? q1 : q0;
  const bool fb = fbits && a == kRoundPeel;
  if (a == kRoundEmpty)
    stamp_all_live(rule.mc, r, thr, state, orig, del_thr);
  else
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x)
      stamp_edge(q[i], r, thr, state, orig, del_thr, src, dst, fb ? fbits : nullptr, ws);
  if (last_block(&ctl->ticket) && threadIdx.x == 0) {
    ctl->alive = a == kRoundEmpty ? 0 : ctl->alive - cnt;
    ctl->rounds_done += 1;
    ctl->stamped = 1;
  }
}

// The frontier of a new level: live edges with support below the threshold
// (the peel kernels run this when the action is kRoundSweep).
__device__ __forceinline__ void sweep_frontier(uint32_t* q0, uint32_t* q1, Ctl* ctl,
                                               const uint32_t* __restrict__ state,
                                               const int32_t* __restrict__ sup, uint32_t mc) {
  const uint32_t cur = ctl->cur;
  uint32_t* q = cur ? q1 : q0;
  const int32_t thr = ctl->thr;
  tile_compact(
      mc, [&](uint64_t e) { return state[e] == kNone && sup[e] < thr; },
      [&](uint32_t at, uint64_t e) { q[at] = uint32_t(e); }, &ctl->qcount[cur]);
}

__device__ __forceinline__ void clear_edge_bits(uint32_t u, uint32_t v, unsigned long long* bits,
                                                unsigned long long* fbits, uint32_t ws) {
  const uint64_t a = uint64_t(u) * ws + (v >> 6), b = uint64_t(v) * ws + (u >> 6);
  const unsigned long long ma = ~(1ull << (v & 63)), mb = ~(1ull << (u & 63));
  atomicAnd(&bits[a], ma);
  atomicAnd(&bits[b], mb);
  if (fbits) {
    atomicAnd(&fbits[a], ma);
    atomicAnd(&fbits[b], mb);
  }
}

// Close a round. After a walked round: clear its frontier from the bit
// matrices; the next frontier is the peel's next queue. After a sweep: the
// next frontier is the swept queue. Every workgroup reads the same frontier
// size, live count and threshold, so all take the same decision on the next
// round and stamp the frontier it walks, drops or empties the level with (a
// compaction takes it unstamped); an empty frontier completes the level
// (level_end). The last workgroup switches queues after a walked round and
// publishes the decision.
__global__ __launch_bounds__(kBlock) void k_round_next(uint32_t* q0, uint32_t* q1, Ctl* ctl, uint32_t* state,
                                                       const int32_t* __restrict__ sup,
                                                       const uint32_t* __restrict__ orig, uint32_t* del_thr,
                                                       const uint32_t* __restrict__ src,
                                                       const uint32_t* __restrict__ dst, unsigned long long* bits,
                                                       unsigned long long* fbits, uint32_t ws, RoundRule rule) {
  const uint32_t a = ctl->action;
  if (a == kRoundMinSup) {
    level_end(ctl, state, sup, rule, [] {});
    return;
  }
  const bool walked = a == kRoundPeel && ctl->stamped;
  if (!walked && a != kRoundSweep) return;
  const uint32_t cur = ctl->cur, r = ctl->round;
  const int32_t thr = ctl->thr;
  const unsigned long long alive = ctl->alive;
  if (walked && bits) {
    const uint32_t cnt = ctl->qcount[cur];
    const uint32_t* q = cur ? q1 : q0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
      const uint32_t e = q[i];
      clear_edge_bits(src[e], dst[e], bits, fbits, ws);
    }
  }
  const uint32_t nq = walked ? cur ^ 1u : cur;  // queue of the next frontier
  const uint32_t c = ctl->qcount[nq];
  auto switch_queue = [&] {
    if (walked) {
      ctl->qcount[cur] = 0;
      ctl->cur = nq;
      ctl->round = r + 1;
    }
  };
  if (c == 0) {
    level_end(ctl, state, sup, rule, switch_queue);
    return;
  }
  const uint32_t d = decide_round(c, alive, thr, rule, walked && !rule.dense);
  const bool stamp = d == kRoundPeel || d == kRoundDrop || d == kRoundEmpty;
  if (stamp) {
    const uint32_t rs = walked ? r + 1 : r;
    const uint32_t* q = nq ? q1 : q0;
    unsigned long long* fb = d == kRoundPeel ? fbits : nullptr;
    if (d == kRoundEmpty)
      stamp_all_live(rule.mc, rs, uint32_t(thr), state, orig, del_thr);
    else
      for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
        stamp_edge(q[i], rs, uint32_t(thr), state, orig, del_thr, src, dst, fb, ws);
  }
  if (last_block(&ctl->ticket) && threadIdx.x == 0) {
    switch_queue();
    if (stamp) {
      ctl->alive = d == kRoundEmpty ? 0 : alive - c;
      ctl->rounds_done += 1;
    }
    ctl->stamped = stamp ? 1u : 0u;
    ctl->action = d;
  }
}

// Clear the current frontier from the bit matrices (a dropped frontier).
__global__ __launch_bounds__(kBlock) void k_frontier_clear(const uint32_t* q0, const uint32_t* q1, const Ctl* ctl,
                                                           const uint32_t* __restrict__ src,
                                                           const uint32_t* __restrict__ dst,
                                                           unsigned long long* bits, unsigned long long* fbits,
                                                           uint32_t ws) {
  const uint32_t cur = ctl->cur, cnt = ctl->qcount[cur];
  const uint32_t* q = cur ? q1 : q0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const uint32_t e = q[i];
    clear_edge_bits(src[e], dst[e], bits, fbits, ws);
  }
}

__device__ __forceinline__ bool dec_crosses(uint32_t x, int32_t thr, int32_t* sup) {
  return x != kNone && atomicSub(&sup[x], 1) == thr;
}

// Walked round on a bit-matrix core. One wave per (frontier edge, chunk of
// its rows); a chunk is 64 words, or 32 or 16 when the frontier is too small
// to give every wave of the grid an item (a short cascade round then spreads
// each edge's triangles over more waves). The common neighbours w of
// e = (u, v) in the chunk are
// the set bits of live[u] & live[v]; their count is scanned across the wave
// and the triangles are dealt out evenly, one per lane per step, whatever
// words they sit in. Frontier bits say at word level whether (u, w) and
// (v, w) are in the frontier, so a triangle costs only the edge-id lookups
// of the partners it decrements. Round-synchronous rule: a triangle with one
// frontier edge charges both partners; with two, the third edge is charged
// once -- by the frontier edge whose non-shared endpoint is larger than the
// apex (z < x for the third edge (x, z)); with three, nobody. The supports
// after the round equal the reference's `2 - #frontier` decrements on
// doubled counts (src/kc.cu:122-146).
template <bool kEidm>
__global__ __launch_bounds__(kBlock) void k_dense_peel_round(
    uint32_t* q0, uint32_t* q1, Ctl* ctl, const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
    const unsigned long long* __restrict__ bits, const unsigned long long* __restrict__ fbits, uint32_t ws,
    const uint32_t* __restrict__ eidm, uint32_t nc, const unsigned long long* __restrict__ bits0,
    const uint32_t* __restrict__ pref0, const uint32_t* __restrict__ off_u, const uint32_t* __restrict__ eid_u,
    const uint32_t* __restrict__ state, int32_t* sup, uint32_t mc, const uint32_t* __restrict__ off_d,
    unsigned long long* stats) {
  const uint32_t action = ctl->action;
  if (action == kRoundSweep) {
    sweep_frontier(q0, q1, ctl, state, sup, mc);
    return;
  }
  if (action != kRoundPeel || !ctl->stamped) return;
  const int32_t thr = ctl->thr;
  const uint32_t cur = ctl->cur;
  const uint32_t cnt = ctl->qcount[cur];
  const uint32_t* q = cur ? q1 : q0;
  uint32_t* qn = cur ? q0 : q1;
  uint32_t* qn_count = &ctl->qcount[cur ^ 1u];
  __shared__ unsigned long long s_t[kBlock], s_fa[kBlock], s_fb[kBlock];
  __shared__ uint32_t s_inc[kBlock];
  const int lane = lane64();
  const uint32_t wbase = threadIdx.x & ~uint32_t(kWave - 1);
  const uint64_t nw = uint64_t(gridDim.x) * (kBlock / kWave);
  uint32_t cw = kWave;  // words per chunk
  while (cw > 16 && uint64_t(cnt) * ((ws + cw - 1) / cw) < nw) cw >>= 1;
  const uint32_t nch = (ws + cw - 1) / cw;
  const uint64_t items = uint64_t(cnt) * nch;
  auto eid_of = [&](uint32_t a, uint32_t w) -> uint32_t {
    if (kEidm) return eidm[uint64_t(a) * nc + w];
    const uint64_t ra = uint64_t(a) * ws;
    const uint32_t i = w >> 6;
    const unsigned long long below = bits0[ra + i] & ((1ull << (w & 63)) - 1ull);
    return eid_u[off_u[a] + pref0[ra + i] + uint32_t(__popcll(below))];
  };
  for (uint64_t it = uint64_t(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; it < items; it += nw) {
    const uint32_t qi = uint32_t(it / nch), ch = uint32_t(it - uint64_t(qi) * nch);
    const uint32_t e = q[qi];
    const uint32_t u = src[e], v = dst[e];
    if (stats && ch == 0 && lane == 0) {  // measurement: the edge's merge work, d+(u) + d+(v)
      atomicAdd(&stats[0], (unsigned long long)(off_d[u + 1] - off_d[u] + off_d[v + 1] - off_d[v]));
      atomicAdd(&stats[1], 1ull);
    }
    const uint64_t ru = uint64_t(u) * ws, rv = uint64_t(v) * ws;
    const uint32_t i = ch * cw + uint32_t(lane);
    unsigned long long t = 0, fa = 0, fb = 0;
    if (uint32_t(lane) < cw && i < ws) {
      t = bits[ru + i] & bits[rv + i];
      fa = fbits[ru + i] & t;
      fb = fbits[rv + i] & t;
      t &= ~(fa & fb);  // both partners in the frontier: nobody is charged
    }
    uint32_t inc = uint32_t(__popcll(t));
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    const uint32_t total = __shfl(inc, kWave - 1);
    s_t[threadIdx.x] = t;
    s_fa[threadIdx.x] = fa;
    s_fb[threadIdx.x] = fb;
    s_inc[threadIdx.x] = inc;
    wave_sync();
    // kPeelUnroll triangles per lane per step: their partner lookups, then
    // all their decrements, are in flight together (the returning atomics
    // are performed past the L2, so one round trip per triangle would leave
    // the wave waiting)
    for (uint32_t j0 = 0; j0 < total; j0 += kWave * kPeelUnroll) {
      uint32_t xa[kPeelUnroll], xb[kPeelUnroll];  // partner edges (u, w) and (v, w) to decrement
#pragma unroll
      for (uint32_t k = 0; k < kPeelUnroll; ++k) {
        xa[k] = kNone;
        xb[k] = kNone;
        const uint32_t j = j0 + k * kWave + uint32_t(lane);
        if (j < total) {
          uint32_t lo = 0, hi = kWave - 1;  // owner: first lane whose inclusive count exceeds j
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_inc[wbase + mid] > j)
              hi = mid;
            else
              lo = mid + 1;
          }
          const unsigned long long tw = s_t[wbase + lo];
          const uint32_t b = select_bit(tw, j - (s_inc[wbase + lo] - uint32_t(__popcll(tw))));
          const uint32_t w = ((ch * cw + lo) << 6) + b;
          const bool ina = (s_fa[wbase + lo] >> b) & 1ull, inb = (s_fb[wbase + lo] >> b) & 1ull;
          if (!ina && !inb) {
            xa[k] = eid_of(u, w);
            xb[k] = eid_of(v, w);
          } else if (ina) {  // (u, w) in the frontier, third edge (v, w)
            if (w < v) xb[k] = eid_of(v, w);
          } else {  // (v, w) in the frontier, third edge (u, w)
            if (w < u) xa[k] = eid_of(u, w);
          }
        }
      }
      int32_t oa[kPeelUnroll], ob[kPeelUnroll];  // supports before the decrements
#pragma unroll
      for (uint32_t k = 0; k < kPeelUnroll; ++k) {
        oa[k] = xa[k] != kNone ? atomicSub(&sup[xa[k]], 1) : 0;
        ob[k] = xb[k] != kNone ? atomicSub(&sup[xb[k]], 1) : 0;
      }
#pragma unroll
      for (uint32_t k = 0; k < kPeelUnroll; ++k) {  // a partner that falls below thr now joins the next frontier
        wave_append(xa[k] != kNone && oa[k] == thr, xa[k], qn, qn_count);
        wave_append(xb[k] != kNone && ob[k] == thr, xb[k], qn, qn_count);
      }
    }
    wave_sync();
  }
}

// Walked round on a sorted-row core: G lanes per frontier edge stream the
// shorter undirected row and binary-search the longer one; partners' states
// say whether they are in this round's frontier (== round) or gone (< round).
template <int G>
__global__ __launch_bounds__(kBlock) void k_sparse_peel_round(
    uint32_t* q0, uint32_t* q1, Ctl* ctl, const uint32_t* __restrict__ off_u, const uint32_t* __restrict__ adj_u,
    const uint32_t* __restrict__ eid_u, const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
    const uint32_t* __restrict__ state, int32_t* sup, uint32_t mc) {
  const uint32_t action = ctl->action;
  if (action == kRoundSweep) {
    sweep_frontier(q0, q1, ctl, state, sup, mc);
    return;
  }
  if (action != kRoundPeel || !ctl->stamped) return;
  const int32_t thr = ctl->thr;
  const uint32_t cur = ctl->cur;
  const uint32_t cnt = ctl->qcount[cur], r = ctl->round;
  const uint32_t* q = cur ? q1 : q0;
  uint32_t* qn = cur ? q0 : q1;
  uint32_t* qn_count = &ctl->qcount[cur ^ 1u];
  const uint32_t groups = gridDim.x * (blockDim.x / G);
  const uint32_t lane = threadIdx.x % G;
  // the longer row of each group's edge is staged in LDS when it fits, so
  // its binary searches cost LDS reads instead of dependent HBM round trips
  constexpr uint32_t kCap = kSparseStage / (kBlock / G);
  __shared__ uint32_t s_rows[kSparseStage];
  uint32_t* sb = s_rows + (threadIdx.x / G) * kCap;
  // every lane of a wave runs the same number of steps so queue appends can
  // be wave-aggregated
  const uint32_t steps = (cnt + groups - 1) / groups;
  for (uint32_t s = 0; s < steps; ++s) {
    const uint32_t qi = s * groups + blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
    uint32_t ab = 0, ae = 0, bb = 0, be = 0, e = 0;
    // A frontier edge with no live triangle (support 0; frontier edges are
    // not decremented within their round) has nobody to charge: no walk.
    if (qi < cnt && sup[q[qi]] > 0) {
      e = q[qi];
      const uint32_t u = src[e], v = dst[e];
      ab = off_u[u], ae = off_u[u + 1];
      bb = off_u[v], be = off_u[v + 1];
      if ((ae - ab) > (be - bb)) {
        uint32_t t = ab; ab = bb; bb = t;
        t = ae; ae = be; be = t;
      }
    }
    // longest shorter-row of the wave bounds the inner loop
    uint32_t len = ae - ab;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) len = max(len, uint32_t(__shfl_xor(len, o)));
    const uint32_t nb = be - bb;
    const bool staged = nb <= kCap;
    wave_sync();  // the previous step's searches are done with the stage
    if (staged && ab < ae)
      for (uint32_t j = lane; j < nb; j += G) sb[j] = adj_u[bb + j];
    wave_sync();
    for (uint32_t i0 = 0; i0 < len; i0 += G) {
      const uint32_t i = ab + i0 + lane;
      uint32_t xa = kNone, xb = kNone;
      if (i < ae) {
        const uint32_t x = adj_u[i];
        const uint32_t p = staged ? bb + lower_bound_u32(sb, 0u, nb, x) : lower_bound_u32(adj_u, bb, be, x);
        if (p < be && (staged ? sb[p - bb] : adj_u[p]) == x) {
          const uint32_t ea = eid_u[i], eb = eid_u[p];
          const uint32_t ma = state[ea], mb = state[eb];
          if (ma >= r && mb >= r) {  // neither side deleted in an earlier round
            const bool ina = (ma == r), inb = (mb == r);
            if (!ina && !inb) {
              xa = ea;
              xb = eb;
            } else if (ina && !inb) {
              if (e < ea) xb = eb;
            } else if (!ina && inb) {
              if (e < eb) xa = ea;
            }
          }
        }
      }
      const bool pa = dec_crosses(xa, thr, sup);
      const bool pb = dec_crosses(xb, thr, sup);
      wave_append(pa, xa, qn, qn_count);
      wave_append(pb, xb, qn, qn_count);
    }
  }
}

// Edge-id matrix of a bit-matrix core: eidm[u * nc + w] = id of edge {u, w}
// (entries of non-edges are never read).
__global__ __launch_bounds__(kBlock) void k_dense_eidm(const uint32_t* __restrict__ off_u,
                                                       const uint32_t* __restrict__ adj_u,
                                                       const uint32_t* __restrict__ eid_u, uint32_t nc,
                                                       uint32_t* eidm) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < nc; u += waves)
    for (uint32_t j = off_u[u] + lane; j < off_u[u + 1]; j += kWave) eidm[uint64_t(u) * nc + adj_u[j]] = eid_u[j];
}

// flag[u] = 1 for every endpoint of a live edge (plain stores of one value).
__global__ void k_live_vertex_flags(const uint32_t* __restrict__ state, const uint32_t* __restrict__ src,
                                    const uint32_t* __restrict__ dst, uint32_t mc, uint32_t* flag) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc; e += gridDim.x * blockDim.x)
    if (state[e] == kNone) {
      flag[src[e]] = 1u;
      flag[dst[e]] = 1u;
    }
}

// Renumber the endpoints of compacted edges through the (monotone) map.
__global__ void k_relabel_edges(uint32_t* adj, uint32_t* src, uint32_t m, const uint32_t* __restrict__ vmap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    adj[i] = vmap[adj[i]];
    src[i] = vmap[src[i]];
  }
}

// Row offsets of the renumbered rows: vertices without a live edge have empty
// rows, so the kept rows stay contiguous.
__global__ void k_relabel_offsets(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ vmap, uint32_t n,
                                  uint32_t* off) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u <= n; u += gridDim.x * blockDim.x)
    if (u == n || vmap[u + 1] != vmap[u]) off[vmap[u]] = roff[u];
}

}  // namespace

// --------------------------------------------------------------- wrappers ---

void live_vertex_flags(const uint32_t* state, const uint32_t* src, const uint32_t* dst, uint32_t mc, uint32_t* flag,
                       hipStream_t s) {
  if (mc) k_live_vertex_flags<<<grid_for(mc), kBlock, 0, s>>>(state, src, dst, mc, flag);
}
void relabel_edges(uint32_t* adj, uint32_t* src, uint32_t m, const uint32_t* vmap, hipStream_t s) {
  if (m) k_relabel_edges<<<grid_for(m), kBlock, 0, s>>>(adj, src, m, vmap);
}
void relabel_offsets(const uint32_t* roff, const uint32_t* vmap, uint32_t n, uint32_t* off, hipStream_t s) {
  k_relabel_offsets<<<grid_for(uint64_t(n) + 1), kBlock, 0, s>>>(roff, vmap, n, off);
}

void round_start(Ctl* ctl, uint32_t r, uint64_t alive, int32_t thr, const RoundRule& rule, bool reset,
                 const SearchRule& search, hipStream_t s) {
  k_round_start<<<1, kWave, 0, s>>>(ctl, r, alive, thr, rule, reset ? 1u : 0u, search);
}

// Workgroups for `items` units of work at `per_block` units each, within
// [1, cap]: a small core launches a few workgroups instead of the full grid
// (every round kernel loops grid-stride, so any grid is correct).
static int work_grid(uint64_t items, uint64_t per_block, int cap) {
  const uint64_t g = (items + per_block - 1) / per_block;
  return int(std::max<uint64_t>(1, std::min<uint64_t>(g, uint64_t(cap))));
}

static void launch_stamp(const PeelArgs& a, Ctl* ctl, const RoundRule& rule, hipStream_t s) {
  k_round_stamp<<<work_grid(a.mc, 4 * kBlock, kStampGrid), kBlock, 0, s>>>(
      a.q0, a.q1, ctl, a.state, a.sup, a.orig, a.del_thr, a.src, a.dst,
      reinterpret_cast<unsigned long long*>(a.fbits), a.ws, rule);
}

void round_stamp(const PeelArgs& a, Ctl* ctl, const RoundRule& rule, hipStream_t s) { launch_stamp(a, ctl, rule, s); }

void peel_rounds(const PeelArgs& a, Ctl* ctl, const RoundRule& rule, int lanes, int count, hipStream_t s,
                 const LaunchProbe& probe) {
  auto bits = reinterpret_cast<unsigned long long*>(a.bits);
  auto fbits = reinterpret_cast<unsigned long long*>(a.fbits);
  auto b0 = reinterpret_cast<const unsigned long long*>(a.bits0);
  static const int next_cap = [] {  // tuning only
    const char* v = std::getenv("KT_NEXT_GRID");
    return v ? std::max(1, std::atoi(v)) : kStampGrid;
  }();
  const int peel_grid = work_grid(a.mc, kBlock, kPeelGrid);
  const int next_grid = work_grid(a.mc, 4 * kBlock, next_cap);
  for (int i = 0; i < count; ++i) {
    if (rule.dense) {
      // every argument as an lvalue for the event-stamped launch
      uint32_t *q0 = a.q0, *q1 = a.q1;
      Ctl* c = ctl;
      const uint32_t *src = a.src, *dst = a.dst, *eidm = a.eidm, *pref0 = a.pref0, *off_u = a.off_u,
                     *eid_u = a.eid_u, *state = a.state, *off_d = a.off_d;
      unsigned long long *lb = bits, *fb = fbits;
      const unsigned long long* b0c = b0;
      uint32_t ws = a.ws, nc = a.nc, mc = a.mc;
      int32_t* sup = a.sup;
      unsigned long long* st = probe.stats;
      void* args[] = {&q0, &q1, &c, &src, &dst, &lb, &fb, &ws, &eidm, &nc, &b0c, &pref0, &off_u, &eid_u,
                      &state, &sup, &mc, &off_d, &st};
      const void* fn = a.eidm ? reinterpret_cast<const void*>(&k_dense_peel_round<true>)
                              : reinterpret_cast<const void*>(&k_dense_peel_round<false>);
      if (probe.ev)
        KT_HIP(hipExtLaunchKernel(fn, dim3(peel_grid), dim3(kBlock), args, 0, s, probe.ev[2 * i],
                                  probe.ev[2 * i + 1], 0));
      else
        KT_HIP(hipLaunchKernel(fn, dim3(peel_grid), dim3(kBlock), args, 0, s));
    } else {
#define KT_SPARSE_ROUND(G)                                                                                 \
  k_sparse_peel_round<G><<<peel_grid, kBlock, 0, s>>>(a.q0, a.q1, ctl, a.off_u, a.adj_u, a.eid_u, a.src, a.dst, \
                                                      a.state, a.sup, a.mc)
      switch (lanes) {
        case 8: KT_SPARSE_ROUND(8); break;
        case 16: KT_SPARSE_ROUND(16); break;
        case 32: KT_SPARSE_ROUND(32); break;
        default: KT_SPARSE_ROUND(64); break;
      }
#undef KT_SPARSE_ROUND
    }
    k_round_next<<<next_grid, kBlock, 0, s>>>(a.q0, a.q1, ctl, a.state, a.sup, a.orig, a.del_thr, a.src, a.dst,
                                              rule.dense ? bits : nullptr, rule.dense ? fbits : nullptr, a.ws,
                                              rule);
  }
}

void frontier_clear(const PeelArgs& a, const Ctl* ctl, hipStream_t s) {
  k_frontier_clear<<<work_grid(a.mc, kBlock, kClearGrid), kBlock, 0, s>>>(a.q0, a.q1, ctl, a.src, a.dst,
                                                 reinterpret_cast<unsigned long long*>(a.bits),
                                                 reinterpret_cast<unsigned long long*>(a.fbits), a.ws);
}

void dense_eidm(const uint32_t* off_u, const uint32_t* adj_u, const uint32_t* eid_u, uint32_t nc, uint32_t* eidm,
                hipStream_t s) {
  if (nc) k_dense_eidm<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(off_u, adj_u, eid_u, nc, eidm);
}

}  // namespace kt
