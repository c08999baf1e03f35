// HIP kernels for the k-max-truss path on CDNA4 (gfx950, wave64).
//
// Stage map (reference counterparts, for parity review):
//   CSR build          src/kc_gpu.cpp:38-65, 243-312 (host parse + init_graph)
//   vertex peel        src/kc.cu:653-884 (__delete_node / kmax_core_gpu)
//   rank + orient      src/kc.cu:1311-1580 (count_deg_with_mask, build_nbr_direct)
//   undirected + ids   src/kc.cu:1583-1673 (build_graph_from_direct)
//   support count      src/kc.cu:193-339, 505-535 (intersection_direct, __tricount_direct)
//   edge peel          src/kc.cu:61-191, 886-1056 (intersection_with_mask, __delete_edge, ktruss)
// The algorithms are restated for wave64 and LDS, not translated: frontier
// queues replace full-array mask scans, hub rows live in LDS as a bitmap plus
// a sorted list, and same-row credits are accumulated in LDS.
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdlib>

#include "kt_device.h"
#include "kt_kernels.h"

namespace kt {
namespace {

// ------------------------------------------------------------ utilities ---

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_iota(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = i;
}
__global__ void k_add(const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = a[i] + b[i];
}
__global__ void k_gather(const uint32_t* idx, const uint32_t* table, uint32_t n, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = table[idx[i]];
}

// ------------------------------------------------------------ (a) CSR ---

__global__ void k_degree_pairs(const uint2* __restrict__ pairs, uint64_t m, uint32_t n,
                               uint32_t* deg, uint32_t* err) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < m;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint2 p = pairs[i];
    if (p.x >= n || p.y >= n) {
      atomicOr(err, 1u);
      continue;
    }
    atomicAdd(&deg[p.x], 1u);
    atomicAdd(&deg[p.y], 1u);
  }
}

__global__ void k_scatter_pairs(const uint2* __restrict__ pairs, uint64_t m, uint32_t n,
                                unsigned long long* cursor, uint32_t* adj) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < m;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint2 p = pairs[i];
    if (p.x >= n || p.y >= n) continue;
    adj[atomicAdd(&cursor[p.x], 1ull)] = p.y;
    adj[atomicAdd(&cursor[p.y], 1ull)] = p.x;
  }
}

__global__ void k_degree_lines(const uint32_t* __restrict__ xs, const uint32_t* __restrict__ ys,
                               uint64_t L, uint32_t n, uint32_t* deg, uint32_t* err) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < L;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t x = xs[i], y = ys[i];
    if (x >= n || y >= n) {
      atomicOr(err, 2u);
      continue;
    }
    atomicAdd(&deg[y], 1u);
  }
}

__global__ void k_scatter_lines(const uint32_t* __restrict__ xs, const uint32_t* __restrict__ ys,
                                uint64_t L, uint32_t n, unsigned long long* cursor, uint32_t* adj) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < L;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t x = xs[i], y = ys[i];
    if (x >= n || y >= n) continue;
    adj[atomicAdd(&cursor[y], 1ull)] = x;
  }
}

// Partitioned CSR build. Directed entries (row, col) are first bucketed by
// row >> S (2^S rows per bucket, small enough for an LDS counter array),
// then each bucket is turned into its rows by one workgroup counting in LDS.
// Entry source: Mode 0 = canonical pairs (both directions), Mode 1 = lines
// (row = ys[i], col = xs[i]).
template <int Mode>
struct EntrySrc {
  const uint32_t* a;
  const uint32_t* b;
  uint64_t items;  // pairs or lines
  __device__ __forceinline__ int per_item() const { return Mode == 0 ? 2 : 1; }
  __device__ __forceinline__ void get(uint64_t i, uint32_t* rows, uint32_t* cols) const {
    if (Mode == 0) {
      const uint2 p = reinterpret_cast<const uint2*>(a)[i];
      rows[0] = p.x;
      cols[0] = p.y;
      rows[1] = p.y;
      cols[1] = p.x;
    } else {
      rows[0] = b[i];
      cols[0] = a[i];
    }
  }
};

// Runs of equal keys in consecutive lanes (sorted inputs put a row's entries
// next to each other): the first lane of each run adds the whole run with
// one LDS atomic instead of one per lane, which would serialise on one
// address. Inactive lanes end runs. head: first lane of this lane's run;
// len: the run's length (meaningful at the head).
struct LaneRun {
  int head;
  uint32_t len;
  bool is_head;
};
__device__ __forceinline__ LaneRun lane_run(uint32_t key, bool active) {
  const int lane = lane64();
  const uint32_t prev = __shfl_up(key, 1);
  const bool prev_active = __shfl_up(int(active), 1) != 0;
  const bool head = active && (lane == 0 || !prev_active || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long ends = heads | ~__ballot(active);  // a run stops at the next head or inactive lane
  const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  LaneRun r;
  r.is_head = head;
  r.head = 63 - __clzll((long long)(heads & le));  // valid for active lanes (lane 0 of a run is a head)
  const unsigned long long after = ends & ~le;
  r.len = uint32_t((after ? __ffsll((unsigned long long)after) - 1 : 64) - lane);
  return r;
}

// kBwd (pairs only): count and scatter only each pair's second entry (row v,
// column u); the first entries of row-sorted pairs are placed in order by
// k_bucket_rows_sorted instead.
template <int Mode, bool kBwd>
__global__ __launch_bounds__(kBlock) void k_part_hist(EntrySrc<Mode> src, uint64_t tile, uint32_t n,
                                                      uint32_t S, uint32_t B, uint32_t* hist_t,
                                                      uint32_t* err) {
  extern __shared__ uint32_t h[];
  const uint32_t t = blockIdx.x;
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint64_t lo = uint64_t(t) * tile;
  const uint64_t hi = min(src.items, lo + tile);
  for (uint64_t i0 = lo; i0 < hi; i0 += blockDim.x) {  // block-uniform trip count: every lane reaches the shuffles
    const uint64_t i = i0 + threadIdx.x;
    uint32_t rows[2] = {0, 0}, cols[2] = {0, 0};
    const bool in = i < hi;
    if (in) src.get(i, rows, cols);
    for (int k = kBwd ? 1 : 0; k < src.per_item(); ++k) {
      bool ok = in;
      if (in && (rows[k] >= n || cols[k] >= n)) {
        atomicOr(err, 1u);
        ok = false;
      }
      const uint32_t b = rows[k] >> S;
      const LaneRun r = lane_run(b, ok);
      if (r.is_head) atomicAdd(&h[b], r.len);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) hist_t[uint64_t(t) * B + i] = h[i];
}

// out[c * rows + r] = in[r * cols + c], through 64 x 64 tiles in LDS (a
// padded row keeps the column reads off one bank).
template <class T>
__global__ __launch_bounds__(kBlock) void k_transpose(const T* __restrict__ in, uint32_t rows, uint32_t cols,
                                                      T* __restrict__ out) {
  __shared__ T tile[64][65];
  const uint32_t c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
  for (uint32_t y = ty; y < 64; y += kBlock / 64) {
    const uint32_t r = r0 + y, c = c0 + tx;
    if (r < rows && c < cols) tile[y][tx] = in[uint64_t(r) * cols + c];
  }
  __syncthreads();
  for (uint32_t y = ty; y < 64; y += kBlock / 64) {
    const uint32_t c = c0 + y, r = r0 + tx;
    if (r < rows && c < cols) out[uint64_t(c) * rows + r] = tile[tx][y];
  }
}

template <int Mode, bool kBwd>
__global__ __launch_bounds__(kBlock) void k_part_scatter(EntrySrc<Mode> src, uint64_t tile, uint32_t n,
                                                         uint32_t S, uint32_t B,
                                                         const uint64_t* __restrict__ hoff_t, uint2* part) {
  extern __shared__ unsigned long long cur[];
  const uint32_t t = blockIdx.x;
  for (uint32_t i = threadIdx.x; i < B; i += blockDim.x) cur[i] = hoff_t[uint64_t(t) * B + i];
  __syncthreads();
  const uint64_t lo = uint64_t(t) * tile;
  const uint64_t hi = min(src.items, lo + tile);
  const int lane = lane64();
  for (uint64_t i0 = lo; i0 < hi; i0 += blockDim.x) {  // block-uniform trip count: every lane reaches the shuffles
    const uint64_t i = i0 + threadIdx.x;
    uint32_t rows[2] = {0, 0}, cols[2] = {0, 0};
    const bool in = i < hi;
    if (in) src.get(i, rows, cols);
    for (int k = kBwd ? 1 : 0; k < src.per_item(); ++k) {
      const bool ok = in && rows[k] < n && cols[k] < n;
      const uint32_t b = rows[k] >> S;
      const LaneRun r = lane_run(b, ok);
      unsigned long long base = 0;
      if (r.is_head) base = atomicAdd(&cur[b], (unsigned long long)r.len);
      base = __shfl(base, ok ? r.head : lane);
      if (ok) part[base + uint64_t(lane - r.head)] = make_uint2(rows[k], cols[k]);
    }
  }
}

// One workgroup per bucket: row counts in LDS, block scan, placement. The
// bucket's entries are part[hoff(b) .. hoff(b + 1)), plus, for row-sorted
// pairs (fwd non-null), the first entries (row u, column v) of the pairs in
// fwd[f0[b] .. f0[b + 1]), which lie there in row order and are read in
// place instead of being scattered; the bucket's rows then start at
// hoff(b) + f0[b]. A row's slots are dealt from one LDS cursor, in any order.
constexpr int kBucketBlock = 1024;
__global__ __launch_bounds__(kBucketBlock) void k_bucket_rows(const uint2* __restrict__ part,
                                                              const uint64_t* __restrict__ hoff,
                                                              const uint2* __restrict__ fwd,
                                                              const uint64_t* __restrict__ f0,
                                                              uint32_t T, uint32_t S, uint32_t n,
                                                              uint64_t* off, uint32_t* deg, uint32_t* adj) {
  extern __shared__ uint32_t cnt[];
  __shared__ uint32_t wsum[kBucketBlock / kWave];
  const uint32_t b = blockIdx.x;
  const uint32_t R = 1u << S;
  const uint32_t r0 = b << S;
  const uint32_t nrows = min(R, n - r0);
  const uint64_t beg = hoff[uint64_t(b) * T], end = hoff[uint64_t(b + 1) * T];
  const uint64_t fb = fwd ? f0[b] : 0, fe = fwd ? f0[b + 1] : 0;
  const uint64_t base = beg + fb;  // the bucket's first slot in adj
  const int lane = lane64();
  for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  auto count = [&](const uint2* __restrict__ src, uint64_t lo, uint64_t hi) {
    for (uint64_t i0 = lo; i0 < hi; i0 += blockDim.x) {  // block-uniform trip count
      const uint64_t i = i0 + threadIdx.x;
      const bool in = i < hi;
      const uint32_t r = in ? src[i].x - r0 : 0u;
      const LaneRun run = lane_run(r, in);
      if (run.is_head) atomicAdd(&cnt[r], run.len);
    }
  };
  count(part, beg, end);
  if (fwd) count(fwd, fb, fe);
  __syncthreads();
  // exclusive scan of cnt[0..R): each thread owns R / blockDim consecutive rows
  const uint32_t per = R / blockDim.x;
  const uint32_t base_i = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t k = 0; k < per; ++k) local += cnt[base_i + k];
  // wave-inclusive scan of `local`
  uint32_t inc = local;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == kWave - 1) wsum[threadIdx.x / kWave] = inc;
  __syncthreads();
  if (threadIdx.x < kWave) {
    uint32_t w = threadIdx.x < blockDim.x / kWave ? wsum[threadIdx.x] : 0;
    uint32_t wi = w;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(wi, o);
      if (int(threadIdx.x) >= o) wi += y;
    }
    if (threadIdx.x < blockDim.x / kWave) wsum[threadIdx.x] = wi - w;
  }
  __syncthreads();
  uint32_t run = wsum[threadIdx.x / kWave] + inc - local;
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t r = base_i + k;
    const uint32_t c = cnt[r];
    if (r < nrows) {
      off[r0 + r] = base + run;
      deg[r0 + r] = c;
    }
    cnt[r] = run;  // becomes the placement cursor
    run += c;
  }
  __syncthreads();
  auto place = [&](const uint2* __restrict__ src, uint64_t lo, uint64_t hi) {
    for (uint64_t i0 = lo; i0 < hi; i0 += blockDim.x) {  // block-uniform trip count
      const uint64_t i = i0 + threadIdx.x;
      const bool in = i < hi;
      const uint2 e = in ? src[i] : make_uint2(r0, 0u);
      const LaneRun run = lane_run(e.x - r0, in);
      uint32_t p = 0;
      if (run.is_head) p = atomicAdd(&cnt[e.x - r0], run.len);
      p = __shfl(p, in ? run.head : lane);
      if (in) adj[base + p + uint32_t(lane - run.head)] = e.y;
    }
  };
  place(part, beg, end);
  if (fwd) place(fwd, fb, fe);
}

// ------------------------------------------- row-sorted inputs (a) ---
// The reference's edge-list contract lists lines sorted by row
// (src/kc_gpu.cpp:243-312 relies on it), and canonical pairs usually come
// sorted by their first id. Such inputs skip the scatter of the lines (their
// rows are runs already) or of the pairs' first entries.

// *unsorted becomes non-zero when keys[i * Stride] > keys[(i + 1) * Stride]
// for some i. With hpos / tpos (lines), each run of equal rows r < n also
// records hpos[r] = its first index and tpos[r] = its last index + 1 (valid
// when the keys are sorted; rows without lines keep what the caller
// filled); a row >= n sets *err.
template <int Stride>
__global__ __launch_bounds__(kBlock) void k_sorted_runs(const uint32_t* __restrict__ keys, uint64_t items,
                                                        uint32_t n, uint64_t* hpos, uint64_t* tpos,
                                                        uint32_t* unsorted, uint32_t* err) {
  bool bad = false, big = false;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < items; i += stride) {
    const uint32_t k = keys[i * Stride];
    const bool has_next = i + 1 < items;
    const uint32_t nx = has_next ? keys[(i + 1) * Stride] : 0u;
    bad |= has_next && nx < k;
    if (hpos) {
      if (k >= n) {
        big = true;
        continue;
      }
      if (i == 0 || keys[(i - 1) * Stride] != k) hpos[k] = i;
      if (!has_next || nx != k) tpos[k] = i + 1;
    }
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(unsorted, 1u);
  if (hpos && __syncthreads_or(big) && threadIdx.x == 0) atomicOr(err, 1u);
}

// deg[r] = tpos[r] - hpos[r] for r < n, deg[n] = 0.
__global__ void k_deg_from_runs(const uint64_t* __restrict__ hpos, const uint64_t* __restrict__ tpos, uint32_t n,
                                uint32_t* deg) {
  for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r <= n; r += uint64_t(gridDim.x) * blockDim.x)
    deg[r] = r < n ? uint32_t(tpos[r] - hpos[r]) : 0u;
}

// adj[i] = xs[i]; a column >= n sets *err.
__global__ __launch_bounds__(kBlock) void k_copy_cols(const uint32_t* __restrict__ xs, uint64_t L, uint32_t n,
                                                      uint32_t* __restrict__ adj, uint32_t* err) {
  bool big = false;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < L; i += stride) {
    const uint32_t x = xs[i];
    big |= x >= n;
    adj[i] = x;
  }
  if (__syncthreads_or(big) && threadIdx.x == 0) atomicOr(err, 1u);
}

// f0[b] = index of the first pair whose first id is >= b << S (pairs sorted
// by their first id), for b < B; f0[B] = items.
__global__ void k_fwd_start(const uint2* __restrict__ pairs, uint64_t items, uint32_t S, uint32_t B,
                            uint64_t* f0) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > B) return;
  uint64_t lo = b < B ? 0 : items, hi = items;
  if (b < B) {
    const uint32_t r0 = b << S;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (pairs[mid].x < r0)
        lo = mid + 1;
      else
        hi = mid;
    }
  }
  f0[b] = lo;
}

__global__ void k_deg_from_off(const uint64_t* off, uint32_t n, uint32_t* deg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    deg[v] = uint32_t(off[v + 1] - off[v]);
}

// ----------------------------------------------------- (b) vertex peel ---

// Device-driven vertex-peel rounds (reference: src/kc.cu:653-884, where each
// round scans a mask and decrements with returned atomics). A round here is
//   sweep  queue and stamp every alive vertex of degree < k (lvl = k);
//   push   every alive neighbour of a queued vertex loses one degree (atomics
//          with no return value: who falls below k is the next sweep's job);
//   close  count the removed vertices; an empty sweep ends the peel.
// The frontier lives in q with ctl->qcount[0] entries; every kernel of a
// round does nothing once the peel has ended (action idle).
__global__ void k_vround_start(Ctl* ctl) {
  if (threadIdx.x != 0) return;
  ctl->qcount[0] = 0;
  ctl->pad[1] = 0;  // heavy-row count
  ctl->removed = 0;
  ctl->action = kRoundPeel;
}

__global__ __launch_bounds__(kBlock) void k_core_sweep(uint32_t* lvl, const uint32_t* __restrict__ deg, uint32_t n,
                                                       uint32_t k, uint32_t* q, Ctl* ctl) {
  if (ctl->action != kRoundPeel) return;
  tile_compact(
      n, [&](uint64_t v) { return lvl[v] == kNone && deg[v] < k; },
      [&](uint32_t at, uint64_t v) {
        q[at] = uint32_t(v);
        lvl[v] = k;
      },
      &ctl->qcount[0]);
}

__global__ void k_vround_close(Ctl* ctl) {
  if (threadIdx.x != 0 || ctl->action != kRoundPeel) return;
  const uint32_t c = ctl->qcount[0];
  ctl->removed += c;
  ctl->qcount[0] = 0;
  ctl->pad[1] = 0;
  if (c == 0) ctl->action = kRoundIdle;
}

// Rows longer than kHeavyRow slots are not walked by one wave: the light
// kernels push them to a heavy list and a persistent kernel then spreads each
// heavy row over all of its blocks (hub rows of power-law graphs would
// otherwise serialise a launch behind one wave).
constexpr uint64_t kHeavyRow = 8192;
constexpr int kHeavyGrid = 512;

__device__ __forceinline__ void push_heavy(uint32_t item, uint32_t* heavy, uint32_t* heavy_count) {
  heavy[atomicAdd(heavy_count, 1u)] = item;
}

// Heavy rows are cut into slices of kSlice slots: slices[h] = slices before
// heavy row h, slices[hc] = total (one block scans the short heavy list).
// The row of heavy item h is row_of(h) = list ? list[heavy[h]] : heavy[h].
constexpr uint64_t kSlice = 4096;
__global__ __launch_bounds__(kBlock) void k_heavy_plan(const uint64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ heavy,
                                                       const uint32_t* heavy_count,
                                                       const uint32_t* __restrict__ list, uint32_t* slices,
                                                       const Ctl* gate) {
  if (gate && gate->action != kRoundPeel) return;
  const uint32_t hc = *heavy_count;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < hc; c0 += blockDim.x) {
    const uint32_t h = c0 + threadIdx.x;
    uint32_t ns = 0;
    if (h < hc) {
      const uint32_t v = list ? list[heavy[h]] : heavy[h];
      ns = uint32_t((off[v + 1] - off[v] + kSlice - 1) / kSlice);
    }
    uint32_t total;
    const uint32_t pre = block_exclusive_scan(ns, &total);
    if (h < hc) slices[h] = carry + pre;
    carry += total;
  }
  if (threadIdx.x == 0) slices[hc] = carry;
}

// slice t -> heavy row h with slices[h] <= t < slices[h + 1]
__device__ __forceinline__ uint32_t slice_row(const uint32_t* slices, uint32_t hc, uint32_t t) {
  uint32_t lo = 0, hi = hc;  // answer in [0, hc)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (slices[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void k_core_push_heavy(const uint64_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj, const uint32_t* heavy,
                                                            const uint32_t* heavy_count,
                                                            const uint32_t* __restrict__ slices,
                                                            const uint32_t* __restrict__ lvl, uint32_t* deg,
                                                            const Ctl* ctl) {
  if (ctl->action != kRoundPeel) return;
  const uint32_t hc = *heavy_count;
  if (hc == 0) return;
  const uint32_t total = slices[hc];
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    const uint32_t h = slice_row(slices, hc, t);
    const uint32_t u = heavy[h];
    const uint64_t b = off[u] + uint64_t(t - slices[h]) * kSlice;
    const uint64_t e = min(off[u + 1], b + kSlice);
    for (uint64_t j = b + threadIdx.x; j < e; j += blockDim.x) {
      const uint32_t w = adj[j];
      if (lvl[w] == kNone) atomicSub(&deg[w], 1u);
    }
  }
}

// Queued vertices are taken in groups of 64: a wave scans the group's row
// lengths and deals the concatenated slots out evenly, so short rows leave no
// lane idle. Each group is shared by P waves (part p takes the 64-slot chunks
// p, p + P, ...), with P sized from the queue so that small frontiers of long
// rows still occupy the whole grid. Rows longer than kHeavyRow go to the
// heavy list for the sliced kernel.
__global__ __launch_bounds__(kBlock) void k_core_push(const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ adj,
                                                      const uint32_t* __restrict__ q, const Ctl* ctl,
                                                      const uint32_t* __restrict__ lvl, uint32_t* deg,
                                                      uint32_t* heavy, uint32_t* heavy_count,
                                                      unsigned long long* stats) {
  if (ctl->action != kRoundPeel) return;
  const uint32_t cnt = ctl->qcount[0];
  __shared__ uint64_t s_b[kBlock];
  __shared__ uint32_t s_inc[kBlock];
  const int lane = lane64();
  const uint32_t wbase = threadIdx.x & ~uint32_t(kWave - 1);
  const uint32_t nw = gridDim.x * (kBlock / kWave);
  const uint32_t groups = (cnt + kWave - 1) / kWave;
  const uint32_t parts = groups ? max(1u, min(uint32_t(kWave), (nw + groups - 1) / groups)) : 1u;
  const uint64_t items = uint64_t(groups) * parts;
  for (uint64_t it = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; it < items; it += nw) {
    const uint32_t grp = uint32_t(it / parts), part = uint32_t(it - uint64_t(grp) * parts);
    const uint32_t i = grp * kWave + uint32_t(lane);
    uint32_t u = 0, len = 0;
    uint64_t b = 0;
    bool hv = false;
    if (i < cnt) {
      u = q[i];
      b = off[u];
      const uint64_t l = off[u + 1] - b;
      hv = l > kHeavyRow;
      len = hv ? 0u : uint32_t(l);
    }
    wave_append<uint32_t>(hv && part == 0, u, heavy, heavy_count);
    uint32_t inc = len;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    const uint32_t total = __shfl(inc, kWave - 1);
    if (stats && part == 0 && lane == 0) {  // measurement: light slots walked, vertices taken
      atomicAdd(&stats[0], (unsigned long long)total);
      atomicAdd(&stats[1], (unsigned long long)min(uint32_t(kWave), cnt - grp * kWave));
    }
    s_b[threadIdx.x] = b;
    s_inc[threadIdx.x] = inc;
    wave_sync();
    for (uint32_t j = part * kWave + uint32_t(lane); j < total; j += kWave * parts) {
      uint32_t lo = 0, hi = kWave - 1;  // owner: first lane whose inclusive length exceeds j
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_inc[wbase + mid] > j)
          hi = mid;
        else
          lo = mid + 1;
      }
      const uint32_t before = lo ? s_inc[wbase + lo - 1] : 0u;
      const uint32_t w = adj[s_b[wbase + lo] + (j - before)];
      if (lvl[w] == kNone) atomicSub(&deg[w], 1u);
    }
    wave_sync();
  }
}

// ------------------------------------------------- vertex-peel cascade ---
//
// One threshold k in one launch. The k-core does not depend on the order in
// which vertices of degree < k are removed, so instead of synchronous rounds
// (a sweep, a push and a close per round, hundreds of rounds on power-law
// graphs) the removals cascade through a work queue: the sweep queues every
// alive vertex of degree < k, and a wave that takes a vertex decrements its
// alive neighbours with returning atomics; the one decrement that takes a
// neighbour's degree from k to k - 1 queues that neighbour (each vertex is
// queued exactly once: by the sweep, or at that crossing). Rows longer than
// kCascadeChunk slots are split into chunk items when taken.
//
// Every neighbour is decremented, removed or not, and lvl is never read: a
// vertex's counter then always equals its number of not-yet-processed
// neighbours, which only falls, and a removed vertex's counter was already
// below its removal threshold <= k, so only an alive vertex can return
// exactly k. (A stamp read from lvl in the same launch could be one left in
// a cache by an earlier launch, so the decision rests on the returning
// atomic alone.)
//
// Queue protocol (counters in Ctl): producers reserve slots with vq_reserve
// and then store their items; a wave claims the next slot with vq_head,
// waits until that slot is reserved and its item stored, processes it, and
// only then adds 1 to vq_done. A wave acts only on values it reads with
// device-coherent atomics (the polled counter or slot itself, or the result
// of a returning atomic), so no acquire or release fence (and no L2
// invalidation or write-back) is needed: every reservation of an item was
// returned before its done increment is issued, so a wave that reads
// vq_done == vq_reserve (done first, reserve after it) sees no item in
// flight, and none can be added any more: every wave leaves there. An item
// past the capacity or a wait past kSpinLimit sets vq_err, which makes every
// wave leave too (the host then fails the solve).
constexpr uint64_t kCascadeChunk = 4096;
constexpr uint32_t kCascadeUnroll = 4;  // groups of 64 row slots per wave step
constexpr unsigned long long kItemEmpty = ~0ull;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t ld_coherent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_coherent(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_item(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_vcascade_start(Ctl* ctl) {
  if (threadIdx.x != 0) return;
  ctl->vq_reserve = 0;
  ctl->vq_head = 0;
  ctl->vq_done = 0;
  ctl->vq_chunks = 0;
  ctl->vq_err = 0;
  ctl->vq_removed = 0;
}

// Initial frontier: every alive vertex of degree < k, stamped lvl = k.
__global__ __launch_bounds__(kBlock) void k_core_sweep_items(uint32_t* lvl, const uint32_t* __restrict__ deg,
                                                             uint32_t n, uint32_t k, unsigned long long* q,
                                                             Ctl* ctl) {
  tile_compact(
      n, [&](uint64_t v) { return lvl[v] == kNone && deg[v] < k; },
      [&](uint32_t at, uint64_t v) {
        q[at] = v;  // chunk 0 of vertex v
        lvl[v] = k;
      },
      &ctl->vq_reserve);
}

// Every wave-uniform value is moved to a scalar register (readfirstlane),
// so the loops below branch on scalars and every lane stays in them; the
// wave leaves through one exit only.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// With stop_at != 0 (the caller restores the level when the k-core is
// empty), a wave that finds stop_at vertices queued -- so fewer than k + 1
// can be left, and a non-empty k-core has at least k + 1 -- raises kVqEmpty,
// and every wave leaves without finishing the cascade.
__global__ void k_vcascade_seed(Ctl* ctl) {
  if (threadIdx.x == 0) ctl->vq_removed = ctl->vq_reserve;  // the sweep queued only vertices
}

__global__ __launch_bounds__(kBlock) void k_core_cascade(const uint64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ adj, uint32_t* lvl,
                                                         uint32_t* deg, uint32_t k, unsigned long long* q,
                                                         uint32_t cap, Ctl* ctl, unsigned long long* stats,
                                                         uint32_t stop_at) {
  const int lane = lane64();
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t quit = 0;
  while (!quit) {
    uint32_t h = 0;
    if (lane == 0) h = atomicAdd(&ctl->vq_head, 1u);
    h = uni(__shfl(h, 0));
    // wait until slot h is reserved (go), or nothing is in flight or an error
    // was raised (quit)
    uint32_t go = 0, spins = 0;
    while (!go && !quit) {
      uint32_t st = 0;
      if (lane == 0) {
        const uint32_t err = ld_coherent(&ctl->vq_err);
        const uint32_t d = ld_coherent(&ctl->vq_done);
        // reserve must be read after done has arrived: the empty asm takes d
        // as an input (the compiler waits for its load) and yields the zero
        // offset of the reserve load, which therefore issues after it
        uint32_t z = 0;
        asm volatile("" : "+v"(z) : "v"(d));
        const uint32_t r = ld_coherent(&ctl->vq_reserve + z);
        st = err ? 2u : (h < r ? 1u : (d == r ? 2u : 0u));
        if (st == 1u && stop_at && ld_coherent(&ctl->vq_removed) >= stop_at) {
          atomicOr(&ctl->vq_err, kVqEmpty);
          st = 2u;
        }
      }
      st = uni(__shfl(st, 0));
      go = st == 1u;
      quit = st == 2u;
      if (st == 0u) {
        if (++spins > kSpinLimit) {
          if (lane == 0) atomicOr(&ctl->vq_err, kVqStall);
          quit = 1;
        } else {
          __builtin_amdgcn_s_sleep(8);
        }
      }
    }
    if (quit) break;
    if (h >= cap) {  // producers never store past cap; a reservation past it is an error
      if (lane == 0) atomicOr(&ctl->vq_err, kVqOverflow);
      break;
    }
    // the slot's producer stores its item right after reserving it
    uint32_t ilo = 0, ihi = 0;
    spins = 0;
    for (;;) {
      uint32_t lo = 0, hi = 0, err = 0;
      if (lane == 0) {
        const unsigned long long x = ld_coherent(&q[h]);
        lo = uint32_t(x);
        hi = uint32_t(x >> 32);
        if (x == kItemEmpty) err = ld_coherent(&ctl->vq_err) & (kVqOverflow | kVqStall);
      }
      ilo = uni(__shfl(lo, 0));
      ihi = uni(__shfl(hi, 0));
      err = uni(__shfl(err, 0));
      if (!(ilo == 0xFFFFFFFFu && ihi == 0xFFFFFFFFu)) break;
      if (err || ++spins > kSpinLimit) {
        if (lane == 0) atomicOr(&ctl->vq_err, kVqStall);
        quit = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (quit) break;
    const uint32_t v = ilo, c = ihi;
    const uint64_t rb = off[v], re = off[v + 1];
    if (c == 0 && re - rb > kCascadeChunk) {  // chunks 1 .. extra of a long row
      const uint32_t extra = uint32_t((re - rb - 1) / kCascadeChunk);
      uint32_t base = 0;
      if (lane == 0) {
        base = atomicAdd(&ctl->vq_reserve, extra);
        atomicAdd(&ctl->vq_chunks, extra);
      }
      base = uni(__shfl(base, 0));
      if (uint64_t(base) + extra > cap) {
        if (lane == 0) atomicOr(&ctl->vq_err, kVqOverflow);
        break;
      }
      for (uint32_t x = uint32_t(lane); x < extra; x += kWave)
        st_item(&q[base + x], (static_cast<unsigned long long>(x + 1) << 32) | v);
    }
    const uint64_t b = rb + uint64_t(c) * kCascadeChunk;
    const uint64_t e = min(re, b + kCascadeChunk);
    // kCascadeUnroll groups of 64 slots per step: their loads, then their
    // returning decrements, are in flight together
    for (uint64_t j0 = b; j0 < e; j0 += kWave * kCascadeUnroll) {
      uint32_t w[kCascadeUnroll];
      bool cross[kCascadeUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kCascadeUnroll; ++u) {
        const uint64_t j = j0 + u * kWave + uint64_t(lane);
        w[u] = j < e ? adj[j] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < kCascadeUnroll; ++u)
        cross[u] = j0 + u * kWave + uint64_t(lane) < e && atomicSub(&deg[w[u]], 1u) == k;
#pragma unroll
      for (uint32_t u = 0; u < kCascadeUnroll; ++u) {
        const unsigned long long m = __ballot(cross[u]);
        if (m) {
          const uint32_t leader = uni(uint32_t(__ffsll(m) - 1));
          const uint32_t cnt = uni(uint32_t(__popcll(m)));
          uint32_t base = 0;
          if (uint32_t(lane) == leader) {
            base = atomicAdd(&ctl->vq_reserve, cnt);
            if (stop_at) atomicAdd(&ctl->vq_removed, cnt);
          }
          base = uni(__shfl(base, int(leader)));
          if (uint64_t(base) + cnt > cap) {
            if (lane == 0) atomicOr(&ctl->vq_err, kVqOverflow);
            quit = 1;
          } else if (cross[u]) {
            lvl[w[u]] = k;
            st_item(&q[base + uint32_t(__popcll(m & below))], static_cast<unsigned long long>(w[u]));
          }
        }
        if (quit) break;
      }
      if (quit) break;
    }
    if (quit) break;
    if (stats && lane == 0) {  // measurement: row slots walked, vertices taken
      atomicAdd(&stats[0], (unsigned long long)(e - b));
      if (c == 0) atomicAdd(&stats[1], 1ull);
    }
    if (lane == 0) atomicAdd(&ctl->vq_done, 1u);
  }
}

__global__ __launch_bounds__(kBlock) void k_core_alive_deg_heavy(const uint64_t* __restrict__ off,
                                                                 const uint32_t* __restrict__ adj,
                                                                 const uint32_t* __restrict__ lvl,
                                                                 const uint32_t* heavy,
                                                                 const uint32_t* heavy_count,
                                                                 const uint32_t* __restrict__ slices,
                                                                 uint32_t* deg) {
  const uint32_t hc = *heavy_count;
  if (hc == 0) return;
  const uint32_t total = slices[hc];
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    const uint32_t h = slice_row(slices, hc, t);
    const uint32_t v = heavy[h];
    const uint64_t b = off[v] + uint64_t(t - slices[h]) * kSlice;
    const uint64_t e = min(off[v + 1], b + kSlice);
    uint32_t c = 0;
    for (uint64_t j = b + threadIdx.x; j < e; j += blockDim.x) c += (lvl[adj[j]] == kNone);
    c = wave_sum(c);
    if (lane64() == 0 && c) atomicAdd(&deg[v], c);
  }
}

__global__ __launch_bounds__(kBlock) void k_core_alive_deg(const uint64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ adj,
                                                           const uint32_t* __restrict__ lvl,
                                                           uint32_t n, uint32_t* deg, uint32_t* heavy,
                                                           uint32_t* heavy_count) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t v = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < n; v += waves) {
    if (lvl[v] != kNone) continue;  // wave-uniform
    if (off[v + 1] - off[v] > kHeavyRow) {
      if (lane == 0) {
        deg[v] = 0;
        push_heavy(v, heavy, heavy_count);
      }
      continue;
    }
    uint32_t c = 0;
    for (uint64_t j = off[v] + lane; j < off[v + 1]; j += kWave) c += (lvl[adj[j]] == kNone);
    c = wave_sum(c);
    if (lane == 0) deg[v] = c;
  }
}

// Degree histogram for the start of the max-core search: exact bins for
// degrees below kDegExact, one bin per power of two above (bin kDegExact + j
// holds degrees in [2^(j + log2 kDegExact), 2^(j + 1 + log2 kDegExact))).
// Counted per workgroup in LDS, merged with one atomic per non-empty bin.
__global__ __launch_bounds__(kBlock) void k_degree_hist(const uint32_t* __restrict__ deg, uint32_t n,
                                                        uint32_t* hist) {
  __shared__ uint32_t s_h[kDegBins];
  for (uint32_t i = threadIdx.x; i < kDegBins; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t d = deg[v];
    const uint32_t b = d < kDegExact ? d : kDegExact + (31u - __clz(d)) - kDegExactLog2;
    atomicAdd(&s_h[b], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kDegBins; i += blockDim.x)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// Vertices of degree < k cannot be in the k-core: stamp them removed at k
// (no decrements: their neighbours' degrees are recounted afterwards) and
// count the rest.
__global__ __launch_bounds__(kBlock) void k_core_filter(const uint32_t* __restrict__ deg, uint32_t n, uint32_t k,
                                                        uint32_t* lvl, uint32_t* alive) {
  uint32_t c = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const bool keep = deg[v] >= k;
    lvl[v] = keep ? kNone : k;
    c += keep;
  }
  c = wave_sum(c);
  if (lane64() == 0 && c) atomicAdd(alive, c);
}

__global__ void k_core_reopen_above(uint32_t* lvl, uint32_t n, uint32_t k) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t l = lvl[v];
    if (l != kNone && l > k) lvl[v] = kNone;
  }
}

// -------------------------------------------- small-level core numbers ---

// The max-core search on a small level in one workgroup: the degrees of the
// level's vertices live in LDS, and thresholds are walked upward from the
// current one without restoring anything (k-cores are nested), each jumping
// to the minimum alive degree + 1. A threshold queues every alive vertex of
// degree < k (lvl = k), and the queue is drained in steps separated by
// barriers, one wave per queued vertex: every neighbour loses one degree, and
// the decrement that takes it from k to k - 1 queues it (an LDS atomic, so
// no vertex is queued twice). When a threshold empties the level, the
// vertices alive at its start are the max core: the ones it removed are
// re-opened (lvl = kNone), and all have the degrees they had when that
// threshold started. A non-empty k-core has at least k + 1 vertices, so a
// threshold is known to empty the level as soon as fewer than k + 1 are left,
// and its cascade stops there (on a dense max core, most of its rows are
// then never walked). A removed vertex's LDS degree carries bit 31;
// afterwards lvl[v] = 1 + the core number of every vertex removed here, so
// every threshold from k0 up is an exact cut.
constexpr uint32_t kDecompBlock = 1024;
constexpr uint32_t kDecompMaxN = 16000;   // LDS degrees: 64,000 B
constexpr uint32_t kDecompUnroll = 8;     // row slots a lane loads before its atomics
constexpr uint32_t kGone = 0x80000000u;
__global__ __launch_bounds__(kDecompBlock) void k_core_decompose(const uint64_t* __restrict__ off,
                                                                  const uint32_t* __restrict__ adj, uint32_t n,
                                                                  uint32_t k0, uint32_t* lvl, uint32_t* deg,
                                                                  uint32_t* q, uint32_t* out) {
  extern __shared__ uint32_t s_deg[];
  __shared__ uint32_t s_tail, s_alive, s_min;
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  constexpr uint32_t kWaves = kDecompBlock / kWave;
  if (tid == 0) {
    s_tail = 0;
    s_alive = 0;
    s_min = kNone;
  }
  __syncthreads();
  for (uint32_t v = tid; v < n; v += kDecompBlock) {
    const bool alive = lvl[v] == kNone;
    s_deg[v] = alive ? deg[v] : kGone;
    if (alive) {
      atomicAdd(&s_alive, 1u);
      atomicMin(&s_min, deg[v]);
    }
  }
  __syncthreads();
  uint32_t k = max(k0, s_min + 1);  // every alive degree is >= s_min: the first removing threshold
  uint32_t start = 0;               // queue position where threshold k's removals begin
  uint32_t alive0 = s_alive;        // vertices alive when threshold k starts
  bool done = alive0 == 0;
  while (!done) {
    // the threshold's first vertices; the alive degrees are kept in deg, so
    // the threshold that empties the level leaves the max core's degrees there
    for (uint32_t v = tid; v < n; v += kDecompBlock) {
      const uint32_t d = s_deg[v];
      if (d & kGone) continue;
      deg[v] = d;
      if (d < k) {
        s_deg[v] = d | kGone;
        lvl[v] = k;
        q[atomicAdd(&s_tail, 1u)] = v;
      }
    }
    __syncthreads();
    // drain: one wave per queued vertex, in steps separated by barriers
    uint32_t head = start;
    bool empties = false;  // fewer than k + 1 vertices left: the k-core is empty
    for (;;) {
      const uint32_t tail = s_tail;
      if (alive0 - (tail - start) < k + 1) {
        empties = true;
        break;
      }
      if (head == tail) break;
      for (uint32_t i = head + wid; i < tail; i += kWaves) {
        const uint32_t v = q[i];
        const uint64_t b = off[v], e = off[v + 1];
        for (uint64_t j0 = b; j0 < e; j0 += kWave * kDecompUnroll) {
          uint32_t w[kDecompUnroll];  // independent loads in flight before the atomics
#pragma unroll
          for (uint32_t u = 0; u < kDecompUnroll; ++u) {
            const uint64_t j = j0 + u * kWave + lane;
            w[u] = j < e ? adj[j] : kNone;
          }
#pragma unroll
          for (uint32_t u = 0; u < kDecompUnroll; ++u)
            if (w[u] != kNone && atomicSub(&s_deg[w[u]], 1u) == k) {  // w falls below k now
              atomicOr(&s_deg[w[u]], kGone);
              lvl[w[u]] = k;
              q[atomicAdd(&s_tail, 1u)] = w[u];
            }
        }
      }
      head = tail;
      __syncthreads();
    }
    if (empties) break;  // s_tail is final: the check above follows a barrier
    // s_tail is final for this threshold
    const uint32_t removed = s_tail - start;
    __syncthreads();
    if (tid == 0) {
      s_alive -= removed;
      s_min = kNone;
    }
    __syncthreads();
    if (s_alive == 0) {
      done = true;
    } else {
      for (uint32_t v = tid; v < n; v += kDecompBlock) {
        const uint32_t d = s_deg[v];
        if (!(d & kGone)) atomicMin(&s_min, d);
      }
      __syncthreads();
      start = s_tail;
      k = s_min + 1;
      alive0 = s_alive;
      __syncthreads();
    }
  }
  // the vertices alive when the last threshold started are the max core:
  // its removals (queue [start, s_tail)) are re-opened, the others were never
  // removed; deg holds all their degrees at that threshold's start, i.e.
  // within the core
  const uint32_t tail = s_tail;
  for (uint32_t i = start + tid; i < tail; i += kDecompBlock) lvl[q[i]] = kNone;
  if (tid == 0) {
    out[0] = alive0 ? k - 1 : 0;  // the max core's k (0: nothing was alive)
    out[1] = alive0;              // its vertices
  }
}

// --------------------------------------------------- (a1-a3) extraction ---

// key = degree-in-core << 32 | (~id): ascending order = cmp_link's
// (degree ascending, id descending), src/kc_gpu.cpp:27-30.
__global__ __launch_bounds__(kBlock) void k_core_collect(const uint32_t* __restrict__ lvl,
                                                         const uint32_t* __restrict__ deg,
                                                         const uint32_t* __restrict__ to_orig, uint32_t n,
                                                         uint64_t* keys, uint32_t* vals, uint32_t* count) {
  tile_compact(
      n, [&](uint64_t v) { return lvl[v] == kNone; },
      [&](uint32_t at, uint64_t v) {
        const uint32_t id = to_orig ? to_orig[v] : uint32_t(v);
        keys[at] = uint64_t(deg[v]) << 32 | uint64_t(0xFFFFFFFFu - id);
        vals[at] = uint32_t(v);
      },
      count);
}

// sorted (key, local id) -> rank maps; inv_orig carries the input's ids
__global__ void k_assign_rank(const uint64_t* keys, const uint32_t* vals, uint32_t nc, uint32_t* to,
                              uint32_t* inv, uint32_t* inv_orig) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nc; r += gridDim.x * blockDim.x) {
    const uint32_t v = vals[r];
    to[v] = r;
    inv[r] = v;
    inv_orig[r] = 0xFFFFFFFFu - uint32_t(keys[r]);
  }
}

// ---------------------------------------------- (b') induced subgraph ---

// Induced subgraph of the alive vertices: sub[i] = alive vertex, new_id[v] = i
// (kNone for the removed vertices, so one lookup tells both).
__global__ __launch_bounds__(kBlock) void k_alive_list(const uint32_t* __restrict__ lvl, uint32_t n,
                                                       uint32_t* list, uint32_t* new_id, uint32_t* count) {
  tile_compact(
      n,
      [&](uint64_t v) {
        const bool alive = lvl[v] == kNone;
        if (!alive) new_id[v] = kNone;
        return alive;
      },
      [&](uint32_t at, uint64_t v) {
        list[at] = uint32_t(v);
        new_id[v] = at;
      },
      count);
}

// Alive vertices as a bitmap (bit v % 32 of word v / 32) with per-word
// prefix counts: the rank of an alive vertex (its id in the cut) is
// pref[word] + the alive bits below it. The two arrays take n / 4 bytes, so
// the rank lookups of a cut stay in L2 and the MALL where an n-entry id map
// would be read from HBM.
// A cut of a large level keeps few vertices, so most lookups find an empty
// word: coarse holds one bit per bitmap word (word w / 32 non-empty), n / 256
// bytes that stay in L2, and only the words it flags are read from bits and
// pref.
__device__ __forceinline__ bool coarse_hit(const uint32_t* __restrict__ coarse, uint32_t w) {
  return (coarse[w >> 10] >> ((w >> 5) & 31)) & 1u;
}

__device__ __forceinline__ uint32_t alive_rank(const uint32_t* __restrict__ bits, const uint32_t* __restrict__ pref,
                                               const uint32_t* __restrict__ coarse, uint32_t w) {
  if (!coarse_hit(coarse, w)) return kNone;
  const uint32_t b = bits[w >> 5], m = 1u << (w & 31);
  return (b & m) ? pref[w >> 5] + uint32_t(__popc(b & (m - 1u))) : kNone;
}

// bits[i] and cnt[i] = popcount for every word; cnt[words] = 0 (the scan's tail).
__global__ __launch_bounds__(kBlock) void k_alive_words(const uint32_t* __restrict__ lvl, uint32_t n, uint32_t* bits,
                                                        uint32_t* cnt) {
  const uint32_t words = (n + 31) / 32;
  const uint32_t groups = (words + 1) / 2;  // 64 vertices per wave step
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t g = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; g < groups; g += waves) {
    const uint32_t v = g * 64 + uint32_t(lane);
    const unsigned long long m = __ballot(v < n && lvl[v] == kNone);
    const uint32_t i = g * 2 + uint32_t(lane);
    if (lane < 2 && i < words) {
      const uint32_t b = uint32_t(m >> (32 * lane));
      bits[i] = b;
      cnt[i] = uint32_t(__popc(b));
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[words] = 0;
}

// coarse[c] bit k = bits[32 c + k] != 0, for c < (words + 31) / 32.
__global__ void k_alive_coarse(const uint32_t* __restrict__ bits, uint32_t words, uint32_t* coarse) {
  const uint32_t cwords = (words + 31) / 32;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cwords; c += gridDim.x * blockDim.x) {
    uint32_t m = 0;
    const uint32_t k1 = min(32u, words - c * 32);
    for (uint32_t k = 0; k < k1; ++k) m |= uint32_t(bits[c * 32 + k] != 0) << k;
    coarse[c] = m;
  }
}

// list[rank] = vertex, ranks in increasing vertex id.
__global__ __launch_bounds__(kBlock) void k_alive_ranked_list(const uint32_t* __restrict__ bits,
                                                              const uint32_t* __restrict__ pref, uint32_t words,
                                                              uint32_t* list) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) {
    uint32_t b = bits[i], r = pref[i];
    while (b) {
      list[r++] = i * 32 + uint32_t(__ffs(b) - 1);
      b &= b - 1u;
    }
  }
}

__global__ void k_sub_degrees(const uint32_t* __restrict__ list, const uint32_t* __restrict__ deg,
                              const uint32_t* __restrict__ to_orig, uint32_t na, uint32_t* sub_deg,
                              uint32_t* sub_orig) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const uint32_t v = list[i];
    if (deg) sub_deg[i] = deg[v];
    sub_orig[i] = to_orig ? to_orig[v] : v;
  }
}

__device__ __forceinline__ uint32_t alive_bit(const uint32_t* __restrict__ bits, const uint32_t* __restrict__ coarse,
                                              uint32_t w) {
  return coarse_hit(coarse, w) ? (bits[w >> 5] >> (w & 31)) & 1u : 0u;
}

// kCutUnroll row slots per lane at once: each level of the lookup (row
// entry, coarse word, fine word, prefix) is loaded for all of them before
// any is used, so their round trips overlap. ok[u] marks the slots in range;
// nid[u] = the neighbour's new id, or kNone when it is not alive (with
// pref == nullptr: 1 / kNone for alive / not).
constexpr uint32_t kCutUnroll = 4;
__device__ __forceinline__ void alive_ranks(const uint32_t* __restrict__ adj, uint64_t j0, uint64_t step,
                                            uint64_t e, const uint32_t* __restrict__ bits,
                                            const uint32_t* __restrict__ pref, const uint32_t* __restrict__ coarse,
                                            uint32_t* nid) {
  uint32_t w[kCutUnroll], c[kCutUnroll], b[kCutUnroll];
#pragma unroll
  for (uint32_t u = 0; u < kCutUnroll; ++u) {
    const uint64_t j = j0 + u * step;
    w[u] = j < e ? adj[j] : kNone;
  }
#pragma unroll
  for (uint32_t u = 0; u < kCutUnroll; ++u) c[u] = w[u] != kNone ? coarse[w[u] >> 10] : 0u;
#pragma unroll
  for (uint32_t u = 0; u < kCutUnroll; ++u) {
    b[u] = 0;
    if ((c[u] >> ((w[u] >> 5) & 31)) & 1u) b[u] = bits[w[u] >> 5];
  }
#pragma unroll
  for (uint32_t u = 0; u < kCutUnroll; ++u) {
    nid[u] = kNone;
    const uint32_t m = 1u << (w[u] & 31);
    if (w[u] != kNone && (b[u] & m)) nid[u] = pref ? pref[w[u] >> 5] + uint32_t(__popc(b[u] & (m - 1u))) : 1u;
  }
}

// Degrees of the cut counted from the rows themselves: sub_deg[i] = the
// neighbours of list[i] whose bit is set in the alive bitmap (n / 8 bytes,
// so the lookups stay in L2 and the MALL). One wave per listed vertex; rows
// longer than kHeavyRow go to the heavy list (by list index) for the sliced
// kernel below.
__global__ __launch_bounds__(kBlock) void k_list_alive_deg(const uint64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ adj,
                                                           const uint32_t* __restrict__ list, uint32_t na,
                                                           const uint32_t* __restrict__ abits,
                                                           const uint32_t* __restrict__ acoarse, uint32_t* sub_deg,
                                                           uint32_t* heavy, uint32_t* heavy_count) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < na; i += waves) {
    const uint32_t v = list[i];
    const uint64_t b = off[v], e = off[v + 1];
    if (e - b > kHeavyRow) {
      if (lane == 0) {
        sub_deg[i] = 0;
        push_heavy(i, heavy, heavy_count);
      }
      continue;
    }
    uint32_t c = 0;
    for (uint64_t j0 = b + lane; j0 < e; j0 += kWave * kCutUnroll) {
      uint32_t nid[kCutUnroll];
      alive_ranks(adj, j0, kWave, e, abits, nullptr, acoarse, nid);
#pragma unroll
      for (uint32_t u = 0; u < kCutUnroll; ++u) c += nid[u] != kNone;
    }
    c = wave_sum(c);
    if (lane == 0) sub_deg[i] = c;
  }
}

__global__ __launch_bounds__(kBlock) void k_list_alive_deg_heavy(const uint64_t* __restrict__ off,
                                                                 const uint32_t* __restrict__ adj,
                                                                 const uint32_t* __restrict__ list,
                                                                 const uint32_t* __restrict__ abits,
                                                                 const uint32_t* __restrict__ acoarse,
                                                                 const uint32_t* heavy,
                                                                 const uint32_t* heavy_count,
                                                                 const uint32_t* __restrict__ slices,
                                                                 uint32_t* sub_deg) {
  const uint32_t hc = *heavy_count;
  if (hc == 0) return;
  const uint32_t total = slices[hc];
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    const uint32_t h = slice_row(slices, hc, t);
    const uint32_t i = heavy[h];
    const uint32_t v = list[i];
    const uint64_t b = off[v] + uint64_t(t - slices[h]) * kSlice;
    const uint64_t e = min(off[v + 1], b + kSlice);
    uint32_t c = 0;
    for (uint64_t j0 = b + threadIdx.x; j0 < e; j0 += uint64_t(blockDim.x) * kCutUnroll) {
      uint32_t nid[kCutUnroll];
      alive_ranks(adj, j0, blockDim.x, e, abits, nullptr, acoarse, nid);
#pragma unroll
      for (uint32_t u = 0; u < kCutUnroll; ++u) c += nid[u] != kNone;
    }
    c = wave_sum(c);
    if (lane64() == 0 && c) atomicAdd(&sub_deg[i], c);
  }
}

// Heavy rows of the cut: all blocks share each row; the output order inside
// a row is free (the vertex peel does not need sorted rows), so positions come
// from a per-row cursor.
__global__ __launch_bounds__(kBlock) void k_sub_fill_heavy(const uint64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ adj,
                                                           const uint32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ abits,
                                                           const uint32_t* __restrict__ apref,
                                                           const uint32_t* __restrict__ acoarse,
                                                           const uint64_t* __restrict__ sub_off,
                                                           uint32_t* sub_adj, const uint32_t* heavy,
                                                           const uint32_t* heavy_count,
                                                           const uint32_t* __restrict__ slices,
                                                           unsigned long long* cursor) {
  const uint32_t hc = *heavy_count;
  if (hc == 0) return;
  const uint32_t total = slices[hc];
  const int lane = lane64();
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    const uint32_t h = slice_row(slices, hc, t);
    const uint32_t i = heavy[h];
    const uint32_t v = list[i];
    const uint64_t b = off[v] + uint64_t(t - slices[h]) * kSlice;
    const uint64_t e = min(off[v + 1], b + kSlice);
    for (uint64_t j0 = b + (threadIdx.x & ~63u); j0 < e; j0 += uint64_t(blockDim.x) * kCutUnroll) {
      uint32_t nid[kCutUnroll];
      alive_ranks(adj, j0 + lane, blockDim.x, e, abits, apref, acoarse, nid);
#pragma unroll
      for (uint32_t u = 0; u < kCutUnroll; ++u) {
        const bool p = nid[u] != kNone;
        const uint64_t mask = __ballot(p);
        if (!mask) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        unsigned long long at = 0;
        if (lane == leader) at = atomicAdd(&cursor[h], (unsigned long long)__popcll(mask));
        at = __shfl(at, leader);
        if (p) sub_adj[sub_off[i] + at + uint64_t(__popcll(mask & ((1ull << lane) - 1ull)))] = nid[u];
      }
    }
  }
}

// sub_deg (nullable): the row's alive-neighbour count, for rows laid out
// at an upper bound of their length (the cut without a counting pass)
__global__ __launch_bounds__(kBlock) void k_sub_fill(const uint64_t* __restrict__ off,
                                                     const uint32_t* __restrict__ adj,
                                                     const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ abits,
                                                     const uint32_t* __restrict__ apref,
                                                     const uint32_t* __restrict__ acoarse, uint32_t na,
                                                     const uint64_t* __restrict__ sub_off,
                                                     uint32_t* sub_adj, uint32_t* heavy,
                                                     uint32_t* heavy_count, uint32_t* sub_deg) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < na; i += waves) {
    const uint32_t v = list[i];
    uint64_t pos = sub_off[i];
    const uint64_t e = off[v + 1];
    if (e - off[v] > kHeavyRow) {
      if (lane == 0) push_heavy(i, heavy, heavy_count);
      continue;
    }
    for (uint64_t j0 = off[v]; j0 < e; j0 += kWave * kCutUnroll) {
      uint32_t nid[kCutUnroll];  // slots j0 + u * 64 + lane, in row order
      alive_ranks(adj, j0 + lane, kWave, e, abits, apref, acoarse, nid);
#pragma unroll
      for (uint32_t u = 0; u < kCutUnroll; ++u) {
        const bool p = nid[u] != kNone;
        const uint64_t mask = __ballot(p);
        if (p) sub_adj[pos + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)))] = nid[u];
        pos += uint32_t(__popcll(mask));
      }
    }
    if (sub_deg && lane == 0) sub_deg[i] = uint32_t(pos - sub_off[i]);
  }
}

// A heavy row's alive-neighbour count is its fill cursor.
__global__ void k_heavy_counts(const uint32_t* heavy, const uint32_t* heavy_count,
                               const unsigned long long* __restrict__ cursor, uint32_t* sub_deg) {
  const uint32_t hc = *heavy_count;
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < hc; h += gridDim.x * blockDim.x)
    sub_deg[heavy[h]] = uint32_t(cursor[h]);
}

// len[i] = the row length of list[i] (an upper bound of its alive part).
__global__ void k_list_row_len(const uint64_t* __restrict__ off, const uint32_t* __restrict__ list, uint32_t na,
                               uint32_t* len) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= na; i += gridDim.x * blockDim.x)
    len[i] = i < na ? uint32_t(off[list[i] + 1] - off[list[i]]) : 0u;
}

// Rows filled at an upper bound of their length, packed: one wave per row
// copies its deg[i] entries from tmp[ub_off[i]] to adj[off[i]].
__global__ __launch_bounds__(kBlock) void k_rows_pack(const uint32_t* __restrict__ tmp,
                                                      const uint64_t* __restrict__ ub_off,
                                                      const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ deg, uint32_t na,
                                                      uint32_t* __restrict__ adj) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < na; i += waves) {
    const uint64_t from = ub_off[i], to = off[i];
    const uint32_t d = deg[i];
    for (uint32_t j = uint32_t(lane); j < d; j += kWave) adj[to + j] = tmp[from + j];
  }
}

// u -> w iff rank(u) < rank(w) (src/kc.cu:1389, 1511).
__global__ __launch_bounds__(kBlock) void k_orient_count(const uint64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ adj,
                                                         const uint32_t* __restrict__ to,
                                                         const uint32_t* __restrict__ inv,
                                                         uint32_t nc, uint32_t* dout) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < nc; r += waves) {
    const uint32_t v = inv[r];
    uint32_t c = 0;
    for (uint64_t j = off[v] + lane; j < off[v + 1]; j += kWave) {
      const uint32_t t = to[adj[j]];
      c += (t != kNone && t > r);
    }
    c = wave_sum(c);
    if (lane == 0) dout[r] = c;
  }
}

__global__ __launch_bounds__(kBlock) void k_orient_fill(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
    const uint32_t* __restrict__ to, const uint32_t* __restrict__ inv, uint32_t nc,
    const uint32_t* __restrict__ off_d, uint32_t* adj_d, uint32_t* src, uint32_t* din) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < nc; r += waves) {
    const uint32_t v = inv[r];
    uint32_t pos = off_d[r];
    const uint64_t e = off[v + 1];
    for (uint64_t j0 = off[v]; j0 < e; j0 += kWave) {
      const uint64_t j = j0 + lane;
      uint32_t t = kNone;
      if (j < e) t = to[adj[j]];
      const bool p = (t != kNone && t > r);
      const uint64_t mask = __ballot(p);
      if (p) {
        const uint32_t at = pos + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
        adj_d[at] = t;
        src[at] = r;
        if (din) atomicAdd(&din[t], 1u);
      }
      pos += uint32_t(__popcll(mask));
    }
  }
}

// Bitonic sort of one row per workgroup in LDS; rows longer than kSortCap
// are skipped (the host routes them to a segmented radix sort).
constexpr uint32_t kSortCap = 8192;

// Rows of 2..64 entries: one wave per row, a bitonic network over the lanes
// (registers and shuffles only; a workgroup per short row would spend most
// of its time in barriers).
__global__ __launch_bounds__(kBlock) void k_sort_rows_wave(const uint32_t* __restrict__ off, uint32_t nrows,
                                                           uint32_t* vals) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; row < nrows; row += waves) {
    const uint32_t b = off[row], len = off[row + 1] - b;
    if (len < 2 || len > uint32_t(kWave)) continue;  // wave-uniform
    uint32_t v = uint32_t(lane) < len ? vals[b + lane] : kNone;
#pragma unroll
    for (int size = 2; size <= kWave; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const uint32_t p = __shfl_xor(v, stride);
        const bool up = (lane & size) == 0, low = (lane & stride) == 0;
        v = (low == up) ? min(v, p) : max(v, p);
      }
    }
    if (uint32_t(lane) < len) vals[b + lane] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_sort_rows(const uint32_t* __restrict__ off,
                                                      uint32_t nrows, uint32_t* vals) {
  __shared__ uint32_t buf[kSortCap];
  for (uint32_t row = blockIdx.x; row < nrows; row += gridDim.x) {
    const uint32_t b = off[row], len = off[row + 1] - b;
    if (len <= uint32_t(kWave) || len > kSortCap) continue;  // block-uniform; short rows: k_sort_rows_wave
    uint32_t P = 2;
    while (P < len) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = (i < len) ? vals[b + i] : kNone;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t t = threadIdx.x; t < P / 2; t += blockDim.x) {
          const uint32_t lo = 2 * t - (t & (stride - 1));
          const uint32_t hi = lo + stride;
          const bool up = ((lo & size) == 0);
          const uint32_t a = buf[lo], c = buf[hi];
          if ((a > c) == up) {
            buf[lo] = c;
            buf[hi] = a;
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) vals[b + i] = buf[i];
    __syncthreads();
  }
}

// Undirected row w = [sources of w's in-edges, ascending] ++ [out(w), ascending]
// with the oriented edge id of every slot (the reference's `fa`).
__global__ __launch_bounds__(kBlock) void k_build_undirected(
    const uint32_t* __restrict__ off_d, const uint32_t* __restrict__ adj_d,
    const uint32_t* __restrict__ din_off, const uint32_t* __restrict__ in_src,
    const uint32_t* __restrict__ in_eid, uint32_t nc, const uint32_t* __restrict__ off_u,
    uint32_t* adj_u, uint32_t* eid_u) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; w < nc; w += waves) {
    const uint32_t ib = din_off[w], in_n = din_off[w + 1] - ib;
    const uint32_t ob = off_d[w], out_n = off_d[w + 1] - ob;
    const uint32_t ub = off_u[w];
    for (uint32_t i = lane; i < in_n; i += kWave) {
      adj_u[ub + i] = in_src[ib + i];
      eid_u[ub + i] = in_eid[ib + i];
    }
    for (uint32_t i = lane; i < out_n; i += kWave) {
      adj_u[ub + in_n + i] = adj_d[ob + i];
      eid_u[ub + in_n + i] = ob + i;
    }
  }
}

// --------------------------------------------------------- (c) support ---

// One workgroup per vertex u (rank order). out(u) is staged in LDS as a
// bitmap over ranks plus a sorted list; sub-groups of G lanes stream out(v)
// for each v in out(u) with coalesced loads and test membership in the
// bitmap. A hit w closes triangle (u, v, w), u < v < w:
//   (u,v) and (u,w) are credited in LDS and flushed once per slot,
//   (v,w) is credited by a global atomic at w's slot in out(v).
template <int G>
__global__ __launch_bounds__(kBlock) void k_support_bitmap(const uint32_t* __restrict__ off_d,
                                                           const uint32_t* __restrict__ adj_d,
                                                           uint32_t nc, uint32_t words,
                                                           uint32_t list_cap, int32_t* sup) {
  extern __shared__ uint32_t lds[];
  uint32_t* bits = lds;
  uint32_t* list = bits + words;
  uint32_t* credit = list + list_cap;
  const uint32_t tid = threadIdx.x;
  for (uint32_t u = blockIdx.x; u < nc; u += gridDim.x) {
    const uint32_t beg = off_d[u];
    const uint32_t du = off_d[u + 1] - beg;
    if (du < 2) continue;  // block-uniform
    for (uint32_t i = tid; i < du; i += blockDim.x) {
      list[i] = adj_d[beg + i];
      credit[i] = 0;
    }
    __syncthreads();
    const uint32_t wlo = list[0] >> 5, whi = list[du - 1] >> 5;
    for (uint32_t i = wlo + tid; i <= whi; i += blockDim.x) bits[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < du; i += blockDim.x) atomicOr(&bits[list[i] >> 5], 1u << (list[i] & 31));
    __syncthreads();
    const uint32_t wmax = list[du - 1];
    const uint32_t sg = tid / G, lane = tid % G, nsg = blockDim.x / G;
    for (uint32_t i = sg; i + 1 < du; i += nsg) {
      const uint32_t v = list[i];
      const uint32_t vb = off_d[v], ve = off_d[v + 1];
      uint32_t cnt = 0;
      for (uint32_t j0 = vb; j0 < ve; j0 += G) {
        if (adj_d[j0] > wmax) break;  // rows are sorted: nothing further can hit
        const uint32_t j = j0 + lane;
        if (j < ve) {
          const uint32_t w = adj_d[j];
          if (w <= wmax && ((bits[w >> 5] >> (w & 31)) & 1u)) {
            ++cnt;
            const uint32_t p = lower_bound_u32(list, i + 1, du, w);
            atomicAdd(&credit[p], 1u);
            atomicAdd(&sup[j], 1);
          }
        }
      }
      cnt = group_sum<G>(cnt);
      if (lane == 0 && cnt) atomicAdd(&credit[i], cnt);
    }
    __syncthreads();
    for (uint32_t i = tid; i < du; i += blockDim.x)
      if (credit[i]) atomicAdd(&sup[beg + i], int32_t(credit[i]));
    __syncthreads();
  }
}

// General fallback: one sub-group per oriented edge e = u -> v. Candidates
// are out(u) after e's slot and out(v); the shorter list is streamed, the
// longer one binary-searched.
template <int G>
__global__ __launch_bounds__(kBlock) void k_support_edges(const uint32_t* __restrict__ off_d,
                                                          const uint32_t* __restrict__ adj_d,
                                                          const uint32_t* __restrict__ src,
                                                          uint32_t mc, int32_t* sup) {
  const uint32_t groups = gridDim.x * (blockDim.x / G);
  const uint32_t lane = threadIdx.x % G;
  for (uint32_t e = blockIdx.x * (blockDim.x / G) + threadIdx.x / G; e < mc; e += groups) {
    const uint32_t u = src[e], v = adj_d[e];
    uint32_t ab = e + 1, ae = off_d[u + 1];
    uint32_t bb = off_d[v], be = off_d[v + 1];
    const bool swap = (ae - ab) > (be - bb);
    if (swap) {
      uint32_t t = ab; ab = bb; bb = t;
      t = ae; ae = be; be = t;
    }
    uint32_t cnt = 0;
    for (uint32_t i0 = ab; i0 < ae; i0 += G) {
      const uint32_t i = i0 + lane;
      if (i < ae) {
        const uint32_t x = adj_d[i];
        const uint32_t p = lower_bound_u32(adj_d, bb, be, x);
        if (p < be && adj_d[p] == x) {
          ++cnt;
          atomicAdd(&sup[i], 1);
          atomicAdd(&sup[p], 1);
        }
      }
    }
    cnt = group_sum<G>(cnt);
    if (lane == 0 && cnt) atomicAdd(&sup[e], int32_t(cnt));
  }
}

// ------------------------------------------------------------ (d) peel ---

__global__ __launch_bounds__(kBlock) void k_edge_frontier(const uint32_t* __restrict__ state,
                                                          const int32_t* __restrict__ sup, uint32_t mc,
                                                          int32_t thr, uint32_t* q, uint32_t* qcount) {
  tile_compact(
      mc, [&](uint64_t e) { return state[e] == kNone && sup[e] < thr; },
      [&](uint32_t at, uint64_t e) { q[at] = uint32_t(e); }, qcount);
}

__global__ __launch_bounds__(kBlock) void k_min_support(const uint32_t* __restrict__ state,
                                                        const int32_t* __restrict__ sup, uint32_t mc,
                                                        uint32_t* out) {
  __shared__ uint32_t s_min[kBlock / kWave];
  uint32_t best = kNone;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc; e += gridDim.x * blockDim.x)
    if (state[e] == kNone) best = min(best, uint32_t(sup[e]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = min(best, uint32_t(__shfl_xor(best, o)));
  if (lane64() == 0) s_min[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / kWave; ++k) best = min(best, s_min[k]);
    if (best != kNone) atomicMin(out, best);
  }
}

// Row index of every slot of a CSR with uint32 offsets.
__global__ __launch_bounds__(kBlock) void k_row_src(const uint32_t* __restrict__ off, uint32_t rows,
                                                    uint32_t* src) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < rows; r += waves)
    for (uint32_t j = off[r] + lane; j < off[r + 1]; j += kWave) src[j] = r;
}

// For every oriented edge e = u -> v (u < v), the slot of e in row u (lson)
// and in row v (rson) of the undirected view (the reference's left_son /
// right_son, src/kc.cu:1656-1666).
__global__ __launch_bounds__(kBlock) void k_sons(const uint32_t* __restrict__ off_u,
                                                 const uint32_t* __restrict__ adj_u,
                                                 const uint32_t* __restrict__ eid_u, uint32_t nc,
                                                 uint32_t* lson, uint32_t* rson) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; w < nc; w += waves)
    for (uint32_t j = off_u[w] + lane; j < off_u[w + 1]; j += kWave) {
      if (w < adj_u[j])
        lson[eid_u[j]] = j;
      else
        rson[eid_u[j]] = j;
    }
}

// ------------------------------------------------ dense cores (bitmaps) ---
// For a core dense enough that a bitmap row (nc bits) is no larger than its
// sorted row, neighbourhoods are rows of a bit matrix: bits0 (all core edges,
// static), bits (edges still alive) and pref0 (popcount of bits0 before each
// word, i.e. the slot of a neighbour in the sorted undirected row).

__global__ __launch_bounds__(kBlock) void k_dense_rows(const uint32_t* __restrict__ off_u,
                                                       const uint32_t* __restrict__ adj_u, uint32_t nc,
                                                       uint32_t ws, unsigned long long* bits0,
                                                       unsigned long long* bits, uint32_t* pref0) {
  extern __shared__ unsigned long long row[];
  __shared__ uint32_t wsum[kBlock / kWave];
  for (uint32_t u = blockIdx.x; u < nc; u += gridDim.x) {
    for (uint32_t i = threadIdx.x; i < ws; i += blockDim.x) row[i] = 0ull;
    __syncthreads();
    for (uint32_t j = off_u[u] + threadIdx.x; j < off_u[u + 1]; j += blockDim.x) {
      const uint32_t w = adj_u[j];
      atomicOr(&row[w >> 6], 1ull << (w & 63));
    }
    __syncthreads();
    const uint64_t base = uint64_t(u) * ws;
    // exclusive prefix of popcounts over the row, in chunks of blockDim words
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < ws; c0 += blockDim.x) {
      const uint32_t i = c0 + threadIdx.x;
      const unsigned long long wv = i < ws ? row[i] : 0ull;
      const uint32_t pc = uint32_t(__popcll(wv));
      uint32_t inc = pc;
      const int lane = lane64();
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
      }
      if (lane == kWave - 1) wsum[threadIdx.x / kWave] = inc;
      __syncthreads();
      uint32_t before = carry;
      for (uint32_t k = 0; k < threadIdx.x / kWave; ++k) before += wsum[k];
      uint32_t total = carry;
      for (uint32_t k = 0; k < blockDim.x / kWave; ++k) total += wsum[k];
      if (i < ws) {
        bits0[base + i] = wv;
        bits[base + i] = wv;
        pref0[base + i] = before + inc - pc;
      }
      carry = total;
      __syncthreads();
    }
  }
}

// sup(e) = |N(u) ∩ N(v)| = popcount(bits[u] & bits[v]) for every live
// oriented edge (all edges when state is null). A workgroup takes one source
// row u at a time (rows grid-stride), stages bits[u] in LDS once, and its
// groups of G lanes walk u's out-edges: each edge then reads one row from
// global memory (bits[v]) instead of two.
constexpr uint32_t kDenseMaxWords = 1024;  // nc <= 65536 on a dense core
template <int G>
__global__ __launch_bounds__(kBlock) void k_dense_support(const uint32_t* __restrict__ off_d,
                                                          const uint32_t* __restrict__ dst, uint32_t nc,
                                                          const unsigned long long* __restrict__ bits,
                                                          uint32_t ws, const uint32_t* __restrict__ state,
                                                          int32_t* sup) {
  __shared__ unsigned long long s_row[kDenseMaxWords];
  const uint32_t grp = threadIdx.x / G, ngrp = blockDim.x / G;
  const uint32_t lane = threadIdx.x % G;
  for (uint32_t u = blockIdx.x; u < nc; u += gridDim.x) {  // workgroup-uniform
    const uint32_t e0 = off_d[u], e1 = off_d[u + 1];
    if (e0 == e1) continue;
    __syncthreads();  // the previous row's readers are done
    const unsigned long long* a = bits + uint64_t(u) * ws;
    for (uint32_t i = threadIdx.x; i < ws; i += blockDim.x) s_row[i] = a[i];
    __syncthreads();
    for (uint32_t e = e0 + grp; e < e1; e += ngrp) {
      if (state && state[e] != kNone) continue;  // group-uniform
      const unsigned long long* b = bits + uint64_t(dst[e]) * ws;
      uint32_t c = 0;
      for (uint32_t i = lane; i < ws; i += G) c += uint32_t(__popcll(s_row[i] & b[i]));
      c = group_sum<G>(c);
      if (lane == 0) sup[e] = int32_t(c);
    }
  }
}

// ---------------------------------------------------- (d3) compaction ---

__global__ __launch_bounds__(kBlock) void k_row_alive_count(const uint32_t* __restrict__ off_d,
                                                            const uint32_t* __restrict__ state,
                                                            uint32_t nc, uint32_t* cnt) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < nc; r += waves) {
    uint32_t c = 0;
    for (uint32_t j = off_d[r] + lane; j < off_d[r + 1]; j += kWave) c += (state[j] == kNone);
    c = wave_sum(c);
    if (lane == 0) cnt[r] = c;
  }
}

// Stable per-row compaction of the live edges (rows stay sorted).
__global__ __launch_bounds__(kBlock) void k_row_compact(
    const uint32_t* __restrict__ off_d, const uint32_t* __restrict__ adj_d,
    const uint32_t* __restrict__ orig, const int32_t* __restrict__ sup,
    const uint32_t* __restrict__ state, uint32_t nc, const uint32_t* __restrict__ off2,
    uint32_t* adj2, uint32_t* src2, uint32_t* orig2, int32_t* sup2) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const int lane = lane64();
  for (uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < nc; r += waves) {
    uint32_t pos = off2[r];
    const uint32_t e = off_d[r + 1];
    for (uint32_t j0 = off_d[r]; j0 < e; j0 += kWave) {
      const uint32_t j = j0 + lane;
      const bool p = j < e && state[j] == kNone;
      const uint64_t mask = __ballot(p);
      if (p) {
        const uint32_t at = pos + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
        adj2[at] = adj_d[j];
        src2[at] = r;
        orig2[at] = orig[j];
        if (sup2) sup2[at] = sup[j];
      }
      pos += uint32_t(__popcll(mask));
    }
  }
}

__global__ void k_count_din(const uint32_t* __restrict__ adj, uint32_t m, uint32_t* din) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    atomicAdd(&din[adj[i]], 1u);
}

// Survivors of the level with threshold `floor` among the base edges.
__global__ void k_truss_pairs(const uint32_t* __restrict__ del_thr, uint32_t m0, int32_t floor,
                              const uint32_t* __restrict__ src0, const uint32_t* __restrict__ dst0,
                              const uint32_t* __restrict__ inv, uint2* out, uint32_t* count) {
  for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < m0;
       base += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t e = base + threadIdx.x;
    bool p = false;
    uint2 pr = make_uint2(0, 0);
    if (e < m0) {
      const uint32_t d = del_thr[e];
      p = (d == kNone) || (int64_t(d) > int64_t(floor));
      if (p) {
        const uint32_t a = inv[src0[e]], b = inv[dst0[e]];
        pr = make_uint2(min(a, b), max(a, b));
      }
    }
    wave_append<uint2>(p, pr, out, count);
  }
}

// Byte-model terms of the last core: W(C+) = sum over oriented edges of
// d+(u) + d+(v), and the same sum over the edges outside the kmax-truss.
__global__ void k_w_stats(const uint32_t* __restrict__ off_d, const uint32_t* __restrict__ src,
                          const uint32_t* __restrict__ dst, const uint32_t* __restrict__ del_thr,
                          uint32_t mc, int32_t floor, unsigned long long* out) {
  unsigned long long w_all = 0, w_out = 0;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < mc; e += gridDim.x * blockDim.x) {
    const uint32_t u = src[e], v = dst[e];
    const unsigned long long w = (off_d[u + 1] - off_d[u]) + (off_d[v + 1] - off_d[v]);
    w_all += w;
    const uint32_t d = del_thr[e];
    if (!(d == kNone || int64_t(d) > int64_t(floor))) w_out += w;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    w_all += __shfl_xor(w_all, o);
    w_out += __shfl_xor(w_out, o);
  }
  if (lane64() == 0) {
    atomicAdd(&out[0], w_all);
    atomicAdd(&out[1], w_out);
  }
}

}  // namespace

// --------------------------------------------------------------- wrappers ---

void w_stats(const uint32_t* off_d, const uint32_t* src, const uint32_t* dst, const uint32_t* del_thr,
             uint32_t mc, int32_t floor, uint64_t* out2, hipStream_t s) {
  if (mc)
    k_w_stats<<<grid_for(mc), kBlock, 0, s>>>(off_d, src, dst, del_thr, mc, floor,
                                              reinterpret_cast<unsigned long long*>(out2));
}
uint32_t dense_words(uint32_t nc) { return (nc + 63) / 64; }
bool dense_fits(uint32_t nc, uint32_t mc) {
  if (nc == 0 || nc > 65536) return false;
  // a bitmap row (ws words of 8 bytes) no larger than two average sorted rows
  return uint64_t(dense_words(nc)) * nc <= 2ull * mc;
}
static int dense_lanes(uint32_t ws) {  // tuning only: KT_DENSE_COUNT_WORDS words per lane
  static const uint32_t per = [] {
    const char* v = std::getenv("KT_DENSE_COUNT_WORDS");
    return v ? uint32_t(std::max(1, std::atoi(v))) : 8u;
  }();
  if (ws >= 64 * per) return 64;
  if (ws >= 32 * per) return 32;
  if (ws >= 16 * per / 2) return 16;
  return 8;
}
void dense_rows(const uint32_t* off_u, const uint32_t* adj_u, uint32_t nc, uint32_t ws, uint64_t* bits0,
                uint64_t* bits, uint32_t* pref0, hipStream_t s) {
  if (!nc) return;
  k_dense_rows<<<nc < 65536 ? nc : 65536, kBlock, sizeof(uint64_t) * ws, s>>>(
      off_u, adj_u, nc, ws, reinterpret_cast<unsigned long long*>(bits0),
      reinterpret_cast<unsigned long long*>(bits), pref0);
}
void dense_support(const uint32_t* off_d, const uint32_t* dst, uint32_t nc, uint32_t mc, const uint64_t* bits,
                   uint32_t ws, const uint32_t* state, int32_t* sup, hipStream_t s) {
  if (!mc || !nc) return;
  if (ws > kDenseMaxWords) fail("dense support: rows longer than the LDS stage");
  auto b = reinterpret_cast<const unsigned long long*>(bits);
  const int grid = int(std::min<uint32_t>(nc, 4096));
  switch (dense_lanes(ws)) {
    case 8: k_dense_support<8><<<grid, kBlock, 0, s>>>(off_d, dst, nc, b, ws, state, sup); break;
    case 16: k_dense_support<16><<<grid, kBlock, 0, s>>>(off_d, dst, nc, b, ws, state, sup); break;
    case 32: k_dense_support<32><<<grid, kBlock, 0, s>>>(off_d, dst, nc, b, ws, state, sup); break;
    default: k_dense_support<64><<<grid, kBlock, 0, s>>>(off_d, dst, nc, b, ws, state, sup); break;
  }
}
void row_src(const uint32_t* off, uint32_t rows, uint32_t* src, hipStream_t s) {
  if (rows) k_row_src<<<grid_for(uint64_t(rows) * kWave), kBlock, 0, s>>>(off, rows, src);
}
void sons(const uint32_t* off_u, const uint32_t* adj_u, const uint32_t* eid_u, uint32_t nc, uint32_t* lson,
          uint32_t* rson, hipStream_t s) {
  if (nc) k_sons<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(off_u, adj_u, eid_u, nc, lson, rson);
}
void row_alive_count(const uint32_t* off_d, const uint32_t* state, uint32_t nc, uint32_t* cnt, hipStream_t s) {
  if (nc) k_row_alive_count<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(off_d, state, nc, cnt);
}
void row_compact(const uint32_t* off_d, const uint32_t* adj_d, const uint32_t* orig, const int32_t* sup,
                 const uint32_t* state, uint32_t nc, const uint32_t* off2, uint32_t* adj2, uint32_t* src2,
                 uint32_t* orig2, int32_t* sup2, hipStream_t s) {
  if (nc)
    k_row_compact<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(off_d, adj_d, orig, sup, state, nc, off2,
                                                                    adj2, src2, orig2, sup2);
}
void count_in_degrees(const uint32_t* adj, uint32_t m, uint32_t* din, hipStream_t s) {
  if (m) k_count_din<<<grid_for(m), kBlock, 0, s>>>(adj, m, din);
}
void truss_pairs(const uint32_t* del_thr, uint32_t m0, int32_t floor, const uint32_t* src0,
                 const uint32_t* dst0, const uint32_t* inv, uint32_t* out, uint32_t* count, hipStream_t s) {
  if (m0)
    k_truss_pairs<<<grid_for(m0), kBlock, 0, s>>>(del_thr, m0, floor, src0, dst0, inv,
                                                 reinterpret_cast<uint2*>(out), count);
}

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n) k_fill_u32<<<grid_for(n), kBlock, 0, s>>>(p, n, v);
}
void fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
  if (n) k_fill_u64<<<grid_for(n), kBlock, 0, s>>>(p, n, v);
}
void iota_u32(uint32_t* p, uint32_t n, hipStream_t s) {
  if (n) k_iota<<<grid_for(n), kBlock, 0, s>>>(p, n);
}
void add_offsets(const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* out, hipStream_t s) {
  if (n) k_add<<<grid_for(n), kBlock, 0, s>>>(a, b, n, out);
}
void gather_u32(const uint32_t* idx, const uint32_t* table, uint32_t n, uint32_t* out, hipStream_t s) {
  if (n) k_gather<<<grid_for(n), kBlock, 0, s>>>(idx, table, n, out);
}

void csr_degree_pairs(const uint32_t* pairs, uint64_t m, uint32_t n, uint32_t* deg, Ctl* ctl,
                      hipStream_t s) {
  if (m) k_degree_pairs<<<grid_for(m), kBlock, 0, s>>>(reinterpret_cast<const uint2*>(pairs), m, n,
                                                      deg, &ctl->err);
}
void csr_degree_lines(const uint32_t* xs, const uint32_t* ys, uint64_t L, uint32_t n, uint32_t* deg,
                      Ctl* ctl, hipStream_t s) {
  if (L) k_degree_lines<<<grid_for(L), kBlock, 0, s>>>(xs, ys, L, n, deg, &ctl->err);
}
void csr_scatter_pairs(const uint32_t* pairs, uint64_t m, uint32_t n, uint64_t* cursor, uint32_t* adj,
                       hipStream_t s) {
  if (m)
    k_scatter_pairs<<<grid_for(m), kBlock, 0, s>>>(reinterpret_cast<const uint2*>(pairs), m, n,
                                                   reinterpret_cast<unsigned long long*>(cursor), adj);
}
void csr_scatter_lines(const uint32_t* xs, const uint32_t* ys, uint64_t L, uint32_t n, uint64_t* cursor,
                       uint32_t* adj, hipStream_t s) {
  if (L)
    k_scatter_lines<<<grid_for(L), kBlock, 0, s>>>(xs, ys, L, n,
                                                   reinterpret_cast<unsigned long long*>(cursor), adj);
}
bool csr_partition_plan(uint32_t n, uint32_t* S, uint32_t* B) {
  // the fewest rows per bucket that keep the buckets at most 4096 (the
  // scatter's cursors, 32 KB of LDS); up to 2^15 rows, whose counters take
  // k_bucket_rows 128 KB of LDS (160 KB per workgroup on gfx950). 4096
  // buckets beat 8192 (2^13 rows) and 12288 (2^12) on the 50M-vertex
  // Chung-Lu graph: the scatter's runs per tile grow with fewer buckets.
  for (uint32_t s = 10; s <= 15; ++s) {
    const uint64_t b = (uint64_t(n) + (1u << s) - 1) >> s;
    if (b <= 4096) {
      *S = s;
      *B = uint32_t(b ? b : 1);
      return true;
    }
  }
  return false;
}
void csr_part_hist(int mode, const uint32_t* a, const uint32_t* b, uint64_t items, uint64_t tile, uint32_t n,
                   uint32_t S, uint32_t B, uint32_t T, uint32_t* hist_t, uint32_t* err, hipStream_t s) {
  const size_t lds = sizeof(uint32_t) * B;
  if (mode == 0)
    k_part_hist<0, false><<<T, kBlock, lds, s>>>(EntrySrc<0>{a, b, items}, tile, n, S, B, hist_t, err);
  else if (mode == 2)
    k_part_hist<0, true><<<T, kBlock, lds, s>>>(EntrySrc<0>{a, b, items}, tile, n, S, B, hist_t, err);
  else
    k_part_hist<1, false><<<T, kBlock, lds, s>>>(EntrySrc<1>{a, b, items}, tile, n, S, B, hist_t, err);
}
void transpose_u32(const uint32_t* in, uint32_t rows, uint32_t cols, uint32_t* out, hipStream_t s) {
  if (rows && cols) k_transpose<uint32_t><<<dim3((cols + 63) / 64, (rows + 63) / 64), kBlock, 0, s>>>(in, rows, cols, out);
}
void transpose_u64(const uint64_t* in, uint32_t rows, uint32_t cols, uint64_t* out, hipStream_t s) {
  if (rows && cols) k_transpose<uint64_t><<<dim3((cols + 63) / 64, (rows + 63) / 64), kBlock, 0, s>>>(in, rows, cols, out);
}
void csr_part_scatter(int mode, const uint32_t* a, const uint32_t* b, uint64_t items, uint64_t tile,
                      uint32_t n, uint32_t S, uint32_t B, uint32_t T, const uint64_t* hoff_t, uint32_t* part,
                      hipStream_t s) {
  const size_t lds = sizeof(unsigned long long) * B;
  uint2* p = reinterpret_cast<uint2*>(part);
  if (mode == 0)
    k_part_scatter<0, false><<<T, kBlock, lds, s>>>(EntrySrc<0>{a, b, items}, tile, n, S, B, hoff_t, p);
  else if (mode == 2)
    k_part_scatter<0, true><<<T, kBlock, lds, s>>>(EntrySrc<0>{a, b, items}, tile, n, S, B, hoff_t, p);
  else
    k_part_scatter<1, false><<<T, kBlock, lds, s>>>(EntrySrc<1>{a, b, items}, tile, n, S, B, hoff_t, p);
}
void csr_bucket_rows(const uint32_t* part, const uint64_t* hoff, const uint32_t* fwd_pairs, const uint64_t* f0,
                     uint32_t B, uint32_t T, uint32_t S, uint32_t n, uint64_t* off, uint32_t* deg, uint32_t* adj,
                     hipStream_t s) {
  const size_t lds = sizeof(uint32_t) << S;
  k_bucket_rows<<<B, kBucketBlock, lds, s>>>(reinterpret_cast<const uint2*>(part), hoff,
                                             reinterpret_cast<const uint2*>(fwd_pairs), f0, T, S, n, off, deg, adj);
}
void csr_sorted_check(int mode, const uint32_t* keys, uint64_t items, uint32_t n, uint64_t* hpos, uint64_t* tpos,
                      uint32_t* unsorted, uint32_t* err, hipStream_t s) {
  if (items < 2 && !hpos) return;
  const int grid = grid_for(items, 4 * kBlock);
  if (mode == 0)
    k_sorted_runs<2><<<grid, kBlock, 0, s>>>(keys, items, n, nullptr, nullptr, unsorted, err);
  else
    k_sorted_runs<1><<<grid, kBlock, 0, s>>>(keys, items, n, hpos, tpos, unsorted, err);
}
void csr_deg_from_runs(const uint64_t* hpos, const uint64_t* tpos, uint32_t n, uint32_t* deg, hipStream_t s) {
  k_deg_from_runs<<<grid_for(uint64_t(n) + 1), kBlock, 0, s>>>(hpos, tpos, n, deg);
}
void csr_copy_cols(const uint32_t* xs, uint64_t L, uint32_t n, uint32_t* adj, uint32_t* err, hipStream_t s) {
  if (L) k_copy_cols<<<grid_for(L, 4 * kBlock), kBlock, 0, s>>>(xs, L, n, adj, err);
}
void csr_fwd_start(const uint32_t* pairs, uint64_t items, uint32_t S, uint32_t B, uint64_t* f0, hipStream_t s) {
  k_fwd_start<<<grid_for(uint64_t(B) + 1), kBlock, 0, s>>>(reinterpret_cast<const uint2*>(pairs), items, S, B, f0);
}
void degrees_from_offsets(const uint64_t* off, uint32_t n, uint32_t* deg, hipStream_t s) {
  if (n) k_deg_from_off<<<grid_for(n), kBlock, 0, s>>>(off, n, deg);
}

void core_round_start(Ctl* ctl, hipStream_t s) { k_vround_start<<<1, kWave, 0, s>>>(ctl); }
void core_round(const uint64_t* off, const uint32_t* adj, uint32_t n, uint32_t* q, Ctl* ctl, uint32_t k,
                uint32_t* lvl, uint32_t* deg, HeavyList hl, unsigned long long* push_stats, hipEvent_t ev0,
                hipEvent_t ev1, hipStream_t s) {
  constexpr int kPushGrid = 2048;
  if (!n) return;
  // grids sized by the vertex count: small graphs launch a few workgroups
  const int push_grid = int(std::min<uint64_t>(kPushGrid, std::max<uint64_t>(4, (uint64_t(n) + 63) / 64)));
  k_core_sweep<<<grid_tiles(n), kBlock, 0, s>>>(lvl, deg, n, k, q, ctl);
  if (ev0 && ev1) {  // measurement: events stamped at the kernel's own start and end
    const uint64_t* a_off = off;
    const uint32_t* a_adj = adj;
    const uint32_t* a_q = q;
    const Ctl* a_ctl = ctl;
    const uint32_t* a_lvl = lvl;
    uint32_t* a_deg = deg;
    uint32_t* a_items = hl.items;
    uint32_t* a_count = hl.count;
    unsigned long long* a_stats = push_stats;
    void* args[] = {&a_off, &a_adj, &a_q, &a_ctl, &a_lvl, &a_deg, &a_items, &a_count, &a_stats};
    KT_HIP(hipExtLaunchKernel(reinterpret_cast<const void*>(&k_core_push), dim3(push_grid), dim3(kBlock), args, 0,
                              s, ev0, ev1, 0));
  } else {
    k_core_push<<<push_grid, kBlock, 0, s>>>(off, adj, q, ctl, lvl, deg, hl.items, hl.count, push_stats);
  }
  if (uint64_t(n) > kHeavyRow + 1) {  // else no row can be heavy
    k_heavy_plan<<<1, kBlock, 0, s>>>(off, hl.items, hl.count, nullptr, hl.slices, ctl);
    k_core_push_heavy<<<kHeavyGrid, kBlock, 0, s>>>(off, adj, hl.items, hl.count, hl.slices, lvl, deg, ctl);
  }
  k_vround_close<<<1, kWave, 0, s>>>(ctl);
}
uint64_t cascade_capacity(uint32_t n, uint64_t slots) { return uint64_t(n) + slots / kCascadeChunk + 2; }
void core_cascade(const uint64_t* off, const uint32_t* adj, uint32_t n, uint32_t k, uint32_t* lvl, uint32_t* deg,
                  unsigned long long* q, uint32_t cap, Ctl* ctl, unsigned long long* stats, hipEvent_t ev0,
                  hipEvent_t ev1, hipStream_t s, uint32_t stop_at) {
  fill_u64(reinterpret_cast<uint64_t*>(q), cap, kItemEmpty, s);
  k_vcascade_start<<<1, kWave, 0, s>>>(ctl);
  if (!n) return;
  k_core_sweep_items<<<grid_tiles(n), kBlock, 0, s>>>(lvl, deg, n, k, q, ctl);
  if (stop_at) k_vcascade_seed<<<1, kWave, 0, s>>>(ctl);
  // enough waves to keep a wide cascade busy; idle ones leave at quiescence
  const int grid = int(std::min<uint64_t>(1024, std::max<uint64_t>(8, (uint64_t(n) + 255) / 256)));
  if (ev0 && ev1) {  // measurement: events stamped at the kernel's own start and end
    const uint64_t* a_off = off;
    const uint32_t* a_adj = adj;
    uint32_t* a_lvl = lvl;
    uint32_t* a_deg = deg;
    uint32_t a_k = k;
    unsigned long long* a_q = q;
    uint32_t a_cap = cap;
    Ctl* a_ctl = ctl;
    unsigned long long* a_stats = stats;
    uint32_t a_stop = stop_at;
    void* args[] = {&a_off, &a_adj, &a_lvl, &a_deg, &a_k, &a_q, &a_cap, &a_ctl, &a_stats, &a_stop};
    KT_HIP(hipExtLaunchKernel(reinterpret_cast<const void*>(&k_core_cascade), dim3(grid), dim3(kBlock), args, 0, s,
                              ev0, ev1, 0));
  } else {
    k_core_cascade<<<grid, kBlock, 0, s>>>(off, adj, lvl, deg, k, q, cap, ctl, stats, stop_at);
  }
}
void core_alive_degrees(const uint64_t* off, const uint32_t* adj, const uint32_t* lvl, uint32_t n,
                        uint32_t* deg, HeavyList hl, hipStream_t s) {
  if (!n) return;
  fill_u32(hl.count, 1, 0, s);
  k_core_alive_deg<<<std::min(grid_for(uint64_t(n) * kWave), 8192), kBlock, 0, s>>>(off, adj, lvl, n, deg,
                                                                                    hl.items, hl.count);
  if (uint64_t(n) > kHeavyRow + 1) {  // else no row can be heavy
    k_heavy_plan<<<1, kBlock, 0, s>>>(off, hl.items, hl.count, nullptr, hl.slices, nullptr);
    k_core_alive_deg_heavy<<<kHeavyGrid, kBlock, 0, s>>>(off, adj, lvl, hl.items, hl.count, hl.slices, deg);
  }
}
void degree_hist(const uint32_t* deg, uint32_t n, uint32_t* hist, hipStream_t s) {
  fill_u32(hist, kDegBins, 0, s);
  if (n) k_degree_hist<<<std::min(grid_for(n), 512), kBlock, 0, s>>>(deg, n, hist);
}
void core_filter(const uint32_t* deg, uint32_t n, uint32_t k, uint32_t* lvl, uint32_t* alive, hipStream_t s) {
  if (n) k_core_filter<<<std::min(grid_for(n), 4096), kBlock, 0, s>>>(deg, n, k, lvl, alive);
}
void core_reopen_above(uint32_t* lvl, uint32_t n, uint32_t k, hipStream_t s) {
  if (n) k_core_reopen_above<<<grid_for(n), kBlock, 0, s>>>(lvl, n, k);
}
uint32_t core_decompose_max_n() { return kDecompMaxN; }
void core_decompose(const uint64_t* off, const uint32_t* adj, uint32_t n, uint32_t k0, uint32_t* lvl, uint32_t* deg,
                    uint32_t* q, uint32_t* out, hipStream_t s) {
  if (n > kDecompMaxN) fail("core_decompose: level too large for the LDS degrees");
  fill_u32(out, 2, 0, s);
  if (n) k_core_decompose<<<1, kDecompBlock, sizeof(uint32_t) * n, s>>>(off, adj, n, k0, lvl, deg, q, out);
}

void core_collect(const uint32_t* lvl, const uint32_t* deg, const uint32_t* to_orig, uint32_t n,
                  uint64_t* keys, uint32_t* vals, uint32_t* count, hipStream_t s) {
  if (n) k_core_collect<<<grid_tiles(n), kBlock, 0, s>>>(lvl, deg, to_orig, n, keys, vals, count);
}
void core_assign_rank(const uint64_t* keys, const uint32_t* vals, uint32_t nc, uint32_t* to, uint32_t* inv,
                      uint32_t* inv_orig, hipStream_t s) {
  if (nc) k_assign_rank<<<grid_for(nc), kBlock, 0, s>>>(keys, vals, nc, to, inv, inv_orig);
}
void alive_list(const uint32_t* lvl, uint32_t n, uint32_t* list, uint32_t* new_id, uint32_t* count,
                hipStream_t s) {
  if (n) k_alive_list<<<grid_tiles(n), kBlock, 0, s>>>(lvl, n, list, new_id, count);
}
void sub_degrees(const uint32_t* list, const uint32_t* deg, const uint32_t* to_orig, uint32_t na,
                 uint32_t* sub_deg, uint32_t* sub_orig, hipStream_t s) {
  if (na) k_sub_degrees<<<grid_for(na), kBlock, 0, s>>>(list, deg, to_orig, na, sub_deg, sub_orig);
}
void list_alive_degrees(const uint64_t* off, const uint32_t* adj, const uint32_t* list, uint32_t na,
                        const uint32_t* abits, const uint32_t* acoarse, uint32_t* sub_deg, HeavyList hl,
                        hipStream_t s) {
  if (!na) return;
  fill_u32(hl.count, 1, 0, s);
  k_list_alive_deg<<<std::min(grid_for(uint64_t(na) * kWave), 8192), kBlock, 0, s>>>(off, adj, list, na, abits,
                                                                                     acoarse, sub_deg, hl.items,
                                                                                     hl.count);
  k_heavy_plan<<<1, kBlock, 0, s>>>(off, hl.items, hl.count, list, hl.slices, nullptr);
  k_list_alive_deg_heavy<<<kHeavyGrid, kBlock, 0, s>>>(off, adj, list, abits, acoarse, hl.items, hl.count,
                                                       hl.slices, sub_deg);
}
void alive_coarse(const uint32_t* bits, uint32_t n, uint32_t* coarse, hipStream_t s) {
  const uint32_t words = (n + 31) / 32;
  if (words) k_alive_coarse<<<grid_for((words + 31) / 32), kBlock, 0, s>>>(bits, words, coarse);
}
void alive_words(const uint32_t* lvl, uint32_t n, uint32_t* bits, uint32_t* cnt, hipStream_t s) {
  const uint32_t words = (n + 31) / 32;
  k_alive_words<<<grid_for(uint64_t((words + 1) / 2) * kWave), kBlock, 0, s>>>(lvl, n, bits, cnt);
}
void alive_ranked_list(const uint32_t* bits, const uint32_t* pref, uint32_t n, uint32_t* list, hipStream_t s) {
  const uint32_t words = (n + 31) / 32;
  if (words) k_alive_ranked_list<<<grid_for(words), kBlock, 0, s>>>(bits, pref, words, list);
}
void sub_fill(const uint64_t* off, const uint32_t* adj, const uint32_t* list, const uint32_t* abits,
              const uint32_t* apref, const uint32_t* acoarse, uint32_t na, const uint64_t* sub_off, uint32_t* sub_adj,
              HeavyList hl, uint64_t* cursor, uint32_t cursor_cap, hipStream_t s, uint32_t* sub_deg) {
  if (!na) return;
  fill_u32(hl.count, 1, 0, s);
  fill_u64(cursor, cursor_cap, 0, s);
  k_sub_fill<<<grid_for(uint64_t(na) * kWave), kBlock, 0, s>>>(off, adj, list, abits, apref, acoarse, na, sub_off,
                                                               sub_adj, hl.items, hl.count, sub_deg);
  k_heavy_plan<<<1, kBlock, 0, s>>>(off, hl.items, hl.count, list, hl.slices, nullptr);
  k_sub_fill_heavy<<<kHeavyGrid, kBlock, 0, s>>>(off, adj, list, abits, apref, acoarse, sub_off, sub_adj, hl.items,
                                                 hl.count, hl.slices, reinterpret_cast<unsigned long long*>(cursor));
  if (sub_deg)
    k_heavy_counts<<<grid_for(cursor_cap), kBlock, 0, s>>>(hl.items, hl.count,
                                                          reinterpret_cast<const unsigned long long*>(cursor), sub_deg);
}
void list_row_lengths(const uint64_t* off, const uint32_t* list, uint32_t na, uint32_t* len, hipStream_t s) {
  k_list_row_len<<<grid_for(uint64_t(na) + 1), kBlock, 0, s>>>(off, list, na, len);
}
void rows_pack(const uint32_t* tmp, const uint64_t* ub_off, const uint64_t* off, const uint32_t* deg, uint32_t na,
               uint32_t* adj, hipStream_t s) {
  if (na) k_rows_pack<<<grid_for(uint64_t(na) * kWave), kBlock, 0, s>>>(tmp, ub_off, off, deg, na, adj);
}
void orient_count(const uint64_t* off, const uint32_t* adj, const uint32_t* to, const uint32_t* inv,
                  uint32_t nc, uint32_t* dout, hipStream_t s) {
  if (nc) k_orient_count<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(off, adj, to, inv, nc, dout);
}
void orient_fill(const uint64_t* off, const uint32_t* adj, const uint32_t* to, const uint32_t* inv,
                 uint32_t nc, const uint32_t* off_d, uint32_t* adj_d, uint32_t* src, uint32_t* din,
                 hipStream_t s) {
  if (nc)
    k_orient_fill<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(off, adj, to, inv, nc, off_d, adj_d,
                                                                    src, din);
}
void sort_rows(const uint32_t* off, uint32_t nrows, uint32_t* vals, uint32_t max_row, hipStream_t s) {
  if (!nrows) return;
  k_sort_rows_wave<<<std::min(grid_for(uint64_t(nrows) * kWave), 16384), kBlock, 0, s>>>(off, nrows, vals);
  if (max_row > uint32_t(kWave)) k_sort_rows<<<nrows < 65536 ? nrows : 65536, kBlock, 0, s>>>(off, nrows, vals);
}
void build_undirected(const uint32_t* off_d, const uint32_t* adj_d, const uint32_t* din_off,
                      const uint32_t* in_src_sorted, const uint32_t* in_eid_sorted, uint32_t nc,
                      const uint32_t* off_u, uint32_t* adj_u, uint32_t* eid_u, hipStream_t s) {
  if (nc)
    k_build_undirected<<<grid_for(uint64_t(nc) * kWave), kBlock, 0, s>>>(
        off_d, adj_d, din_off, in_src_sorted, in_eid_sorted, nc, off_u, adj_u, eid_u);
}

static uint32_t bitmap_lds_bytes(uint32_t nc, uint32_t max_dout) {
  const uint32_t words = (nc + 31) / 32;
  const uint32_t cap = (max_dout + 63) & ~63u;
  return 4u * (words + 2u * cap);
}
bool support_bitmap_fits(uint32_t nc, uint32_t max_dout) {
  return bitmap_lds_bytes(nc, max_dout) <= 64u * 1024u;
}
void support_bitmap(const uint32_t* off_d, const uint32_t* adj_d, uint32_t nc, uint32_t max_dout,
                    int lanes, int32_t* sup, Ctl* ctl, hipStream_t s) {
  (void)ctl;
  if (!nc) return;
  const uint32_t words = (nc + 31) / 32;
  const uint32_t cap = (max_dout + 63) & ~63u;
  const uint32_t lds = bitmap_lds_bytes(nc, max_dout);
  const int grid = int(nc < (1u << 20) ? nc : (1u << 20));
  switch (lanes) {
    case 8: k_support_bitmap<8><<<grid, kBlock, lds, s>>>(off_d, adj_d, nc, words, cap, sup); break;
    case 16: k_support_bitmap<16><<<grid, kBlock, lds, s>>>(off_d, adj_d, nc, words, cap, sup); break;
    case 32: k_support_bitmap<32><<<grid, kBlock, lds, s>>>(off_d, adj_d, nc, words, cap, sup); break;
    default: k_support_bitmap<64><<<grid, kBlock, lds, s>>>(off_d, adj_d, nc, words, cap, sup); break;
  }
}
void support_edges(const uint32_t* off_d, const uint32_t* adj_d, const uint32_t* src, uint32_t nc,
                   uint32_t mc, int lanes, int32_t* sup, hipStream_t s) {
  (void)nc;
  if (!mc) return;
  switch (lanes) {
    case 8: k_support_edges<8><<<grid_for(uint64_t(mc) * 8), kBlock, 0, s>>>(off_d, adj_d, src, mc, sup); break;
    case 16: k_support_edges<16><<<grid_for(uint64_t(mc) * 16), kBlock, 0, s>>>(off_d, adj_d, src, mc, sup); break;
    case 32: k_support_edges<32><<<grid_for(uint64_t(mc) * 32), kBlock, 0, s>>>(off_d, adj_d, src, mc, sup); break;
    default: k_support_edges<64><<<grid_for(uint64_t(mc) * 64), kBlock, 0, s>>>(off_d, adj_d, src, mc, sup); break;
  }
}

void edge_frontier_init(const uint32_t* state, const int32_t* sup, uint32_t mc, int32_t thr, uint32_t* q,
                        uint32_t* qcount, hipStream_t s) {
  if (mc) k_edge_frontier<<<grid_tiles(mc), kBlock, 0, s>>>(state, sup, mc, thr, q, qcount);
}
void edge_min_support(const uint32_t* state, const int32_t* sup, uint32_t mc, uint32_t* out, hipStream_t s) {
  if (mc) k_min_support<<<std::min(grid_for(mc), 2048), kBlock, 0, s>>>(state, sup, mc, out);
}

}  // namespace kt
