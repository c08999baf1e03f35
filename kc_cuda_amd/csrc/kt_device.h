// Device-side helpers shared by the kernel translation units (wave64).
#pragma once
#include "kt_common.h"

namespace kt {
namespace {

constexpr int kMaxGrid = 1 << 20;

inline int grid_for(uint64_t items, int per_block = kBlock) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > uint64_t(kMaxGrid)) g = kMaxGrid;
  return int(g);
}

__device__ __forceinline__ int lane64() { return int(__lane_id()); }

// Orders LDS writes and reads between the lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-aggregated queue append. Must be reached by every lane of the wave.
template <typename T>
__device__ __forceinline__ void wave_append(bool pred, T item, T* q, uint32_t* count) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int lane = lane64();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, uint32_t(__popcll(mask)));
  base = __shfl(base, leader);
  if (pred) q[base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)))] = item;
}

// Block-wide exclusive scan of one value per thread (blockDim == kBlock).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t s_w[kBlock / kWave];
  const int lane = lane64(), wid = int(threadIdx.x) / kWave;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == kWave - 1) s_w[wid] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBlock / kWave; ++k) {
    const uint32_t w = s_w[k];
    if (k < wid) before += w;
    tot += w;
  }
  __syncthreads();
  *total = tot;
  return before + inc - v;
}

// Stream compaction of [0, n) by tiles of kBlock * kTileItems indices: one
// counter atomic per tile instead of one per wave (a single queue counter hit
// by every wave of a 4M-vertex sweep serialises the whole launch).
constexpr int kTileItems = 16;
template <typename Pred, typename Emit>
__device__ __forceinline__ void tile_compact(uint64_t n, Pred pred, Emit emit, uint32_t* count) {
  __shared__ uint32_t s_base;
  const uint64_t tile = uint64_t(blockDim.x) * kTileItems;
  for (uint64_t t0 = uint64_t(blockIdx.x) * tile; t0 < n; t0 += uint64_t(gridDim.x) * tile) {
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const uint64_t i = t0 + uint64_t(k) * blockDim.x + threadIdx.x;
      if (i < n && pred(i)) bits |= 1u << k;
    }
    uint32_t total;
    const uint32_t pre = block_exclusive_scan(uint32_t(__popc(bits)), &total);
    if (threadIdx.x == 0 && total) s_base = atomicAdd(count, total);
    __syncthreads();
    uint32_t at = s_base + pre;
#pragma unroll
    for (int k = 0; k < kTileItems; ++k)
      if ((bits >> k) & 1u) emit(at++, t0 + uint64_t(k) * blockDim.x + threadIdx.x);
    __syncthreads();
  }
}

inline int grid_tiles(uint64_t n) {
  const uint64_t tile = uint64_t(kBlock) * kTileItems;
  uint64_t g = (n + tile - 1) / tile;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return int(g);
}

template <int G>
__device__ __forceinline__ uint32_t group_sum(uint32_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return group_sum<64>(v); }

// first index in [lo, hi) with a[idx] >= x
template <typename P>
__device__ __forceinline__ uint32_t lower_bound_u32(const P a, uint32_t lo, uint32_t hi, uint32_t x) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Position of the k-th (0-based) set bit of x; k < popcount(x).
__host__ __device__ __forceinline__ uint32_t select_bit(unsigned long long x, uint32_t k) {
  uint32_t pos = 0, c;
  c = uint32_t(__builtin_popcount(uint32_t(x)));
  if (k >= c) { k -= c; x >>= 32; pos += 32; }
  c = uint32_t(__builtin_popcount(uint32_t(x) & 0xFFFFu));
  if (k >= c) { k -= c; x >>= 16; pos += 16; }
  c = uint32_t(__builtin_popcount(uint32_t(x) & 0xFFu));
  if (k >= c) { k -= c; x >>= 8; pos += 8; }
  c = uint32_t(__builtin_popcount(uint32_t(x) & 0xFu));
  if (k >= c) { k -= c; x >>= 4; pos += 4; }
  c = uint32_t(__builtin_popcount(uint32_t(x) & 0x3u));
  if (k >= c) { k -= c; x >>= 2; pos += 2; }
  if (k >= (uint32_t(x) & 1u)) pos += 1;
  return pos;
}

// Last workgroup of a launch to arrive here (every workgroup must call it
// exactly once, with all its threads). The counter is reset by the winner.
// No fence: an agent-scope fence writes back and invalidates the whole L2 of
// the workgroup's XCD on this architecture, which cost every round kernel
// its cached rows. What a last workgroup needs from the others travels
// through device atomics, which are performed past the L2: each workgroup
// waits for its own (returning) atomics before taking its ticket, and the
// last one reads the results with coherent loads (ld_agent). Plain stores
// are for later launches, which the launch boundary orders.
__device__ __forceinline__ bool last_block(uint32_t* ticket) {
  __shared__ bool s_last;
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (s_last && threadIdx.x == 0) atomicExch(ticket, 0u);
  return s_last;
}

// Load that reads the device-coherent value (past this XCD's L2).
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace
}  // namespace kt
