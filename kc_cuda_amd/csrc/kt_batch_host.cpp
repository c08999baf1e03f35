// Host side of the batched solve: plan (processing order, lanes per graph),
// scratch slabs, autotuner and hipGraph capture of the persistent launch.
#include "kt_batch_host.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace kt {

namespace {
uint32_t bucket_of(uint64_t m) {
  uint32_t b = 0;
  while (b < 40 && (uint64_t(1) << (b + 1)) <= m) ++b;
  return b;
}
}  // namespace

BatchRunner::BatchRunner(Solver* s) : stream_(s->stream()), device_(s->device()) {
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_));
  std::fill(std::begin(table_), std::end(table_), uint8_t(0));
  if (const char* v = std::getenv("KT_BATCH_PROFILE")) profile_ = v[0] == '1';
}

BatchRunner::~BatchRunner() {
  (void)hipSetDevice(device_);
  (void)hipStreamSynchronize(stream_);
  drop_graph();
  DevBuf* all[] = {&poff_, &ns_, &order_, &lanes_, &slab_, &work_, &ticks_, &phases_};
  for (auto* b : all) b->release();
}

void BatchRunner::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

// Lanes per edge for a graph: the tuned choice of its size bucket, else by
// average degree (a group of lanes walks one neighbourhood).
uint8_t BatchRunner::lanes_for(uint64_t m, uint32_t n) const {
  const uint8_t t = table_[bucket_of(m)];
  if (t) return t;
  const double d = n ? 2.0 * double(m) / double(n) : 0.0;
  if (d <= 8) return 4;
  if (d <= 16) return 8;
  if (d <= 48) return 16;
  return 32;
}

BatchArgs BatchRunner::plan(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                            const std::vector<uint32_t>& order, const std::vector<uint8_t>& lanes, int* grid) {
  uint32_t max_n = 0;
  uint64_t max_m = 0;
  for (uint32_t g : order) {
    max_n = std::max(max_n, h_ns[g]);
    max_m = std::max<uint64_t>(max_m, h_off[g + 1] - h_off[g]);
  }
  if (max_m >= (uint64_t(1) << 31)) fail("batch: a graph has 2^31 or more edges");
  for (uint32_t g = 0; g < count; ++g)
    if (h_off[g + 1] < h_off[g]) fail("batch: offsets must be non-decreasing");
  const size_t G = order.size();
  poff_.ensure(sizeof(uint64_t) * (size_t(count) + 1));
  ns_.ensure(sizeof(uint32_t) * std::max<size_t>(count, 1));
  order_.ensure(sizeof(uint32_t) * std::max<size_t>(G, 1));
  lanes_.ensure(std::max<size_t>(count, 1));
  work_.ensure(sizeof(uint32_t) * 4);
  KT_HIP(hipMemcpyAsync(poff_.p, h_off, sizeof(uint64_t) * (size_t(count) + 1), hipMemcpyHostToDevice, stream_));
  if (count) KT_HIP(hipMemcpyAsync(ns_.p, h_ns, sizeof(uint32_t) * count, hipMemcpyHostToDevice, stream_));
  if (G) KT_HIP(hipMemcpyAsync(order_.p, order.data(), sizeof(uint32_t) * G, hipMemcpyHostToDevice, stream_));
  if (count) KT_HIP(hipMemcpyAsync(lanes_.p, lanes.data(), count, hipMemcpyHostToDevice, stream_));
  *grid = int(std::max<size_t>(1, std::min<size_t>(G, size_t(4) * size_t(cus_))));
  const uint64_t words = batch_slab_words(max_n, max_m);
  slab_.ensure(sizeof(uint32_t) * words * size_t(*grid));
  BatchArgs a{};
  a.pairs = d_pairs;
  a.poff = poff_.as<uint64_t>();
  a.ns = ns_.as<uint32_t>();
  a.order = order_.as<uint32_t>();
  a.lanes = lanes_.as<uint8_t>();
  a.G = uint32_t(G);
  a.slab = slab_.as<uint32_t>();
  a.slab_words = words;
  a.work = work_.as<uint32_t>();
  return a;
}

void BatchRunner::solve(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                        uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep, bool use_graph) {
  KT_HIP(hipSetDevice(device_));
  // largest graphs first: long solves start early, short ones fill the tail
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return h_off[a + 1] - h_off[a] > h_off[b + 1] - h_off[b];
  });
  std::vector<uint8_t> lanes(count);
  for (uint32_t g = 0; g < count; ++g) lanes[g] = lanes_for(h_off[g + 1] - h_off[g], h_ns[g]);
  int grid = 0;
  BatchArgs a = plan(d_pairs, h_off, h_ns, count, order, lanes, &grid);
  a.kmax = d_kmax;
  a.count = reinterpret_cast<unsigned long long*>(d_count);
  a.status = d_status;
  a.keep = d_keep;
  a.ticks = nullptr;
  if (!count) return;
  if (profile_) {
    phases_.ensure(sizeof(unsigned long long) * kBatchPhases * count);
    KT_HIP(hipMemsetAsync(phases_.p, 0, sizeof(unsigned long long) * kBatchPhases * count, stream_));
    a.phases = phases_.as<unsigned long long>();
  }
  if (!use_graph) {
    batch_launch(a, grid, stream_);
    KT_HIP(hipStreamSynchronize(stream_));
    if (profile_) report_phases(h_off, count);
    return;
  }
  if (!exec_ || grid != cached_grid_ || std::memcmp(&a, &cached_, sizeof(a)) != 0) {
    drop_graph();
    KT_HIP(hipStreamSynchronize(stream_));
    KT_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    batch_launch(a, grid, stream_);
    KT_HIP(hipStreamEndCapture(stream_, &graph_));
    KT_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    cached_ = a;
    cached_grid_ = grid;
  }
  KT_HIP(hipGraphLaunch(exec_, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  if (profile_) report_phases(h_off, count);
}

// Sums of the per-graph phase times (ms) and mean levels / rounds, by size
// bucket floor(log2 m).
void BatchRunner::report_phases(const uint64_t* h_off, uint32_t count) {
  int khz = 0;
  KT_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
  const double ms_per_tick = khz > 0 ? 1.0 / double(khz) : 1e-5;
  std::vector<unsigned long long> ph(size_t(kBatchPhases) * count);
  KT_HIP(hipMemcpy(ph.data(), phases_.p, sizeof(unsigned long long) * ph.size(), hipMemcpyDeviceToHost));
  struct Acc {
    uint32_t graphs = 0;
    double ms[6] = {0, 0, 0, 0, 0, 0};
    double lv[2] = {0, 0}, rd[2] = {0, 0};
    double max_ms = 0;
  };
  std::vector<Acc> acc(41);
  unsigned long long t_first = ~0ull, t_last_start = 0, t_end = 0, longest = 0;
  uint32_t g_longest = 0;
  double busy = 0;
  for (uint32_t g = 0; g < count; ++g) {
    const unsigned long long* p = &ph[size_t(g) * kBatchPhases];
    Acc& a = acc[bucket_of(h_off[g + 1] - h_off[g])];
    a.graphs += 1;
    t_first = std::min(t_first, p[8]);
    t_last_start = std::max(t_last_start, p[8]);
    t_end = std::max(t_end, p[9]);
    busy += double(p[9] - p[8]);
    a.max_ms = std::max(a.max_ms, double(p[9] - p[8]) * ms_per_tick);
    if (p[9] - p[8] > longest) {
      longest = p[9] - p[8];
      g_longest = g;
    }
    unsigned long long prev = 0;
    for (int i = 0; i < 6; ++i) {
      const unsigned long long t = std::max(p[i], prev);  // a skipped pass leaves 0
      a.ms[i] += double(t - prev) * ms_per_tick;
      prev = t;
    }
    for (int k = 0; k < 2; ++k) {
      a.lv[k] += double(p[6 + k] >> 32);
      a.rd[k] += double(p[6 + k] & 0xffffffffu);
    }
  }
  std::fprintf(stderr,
               "[batch] span %.1f ms, last start at %.1f ms, busy %.1f workgroup-s (mean concurrency %.0f); "
               "longest graph %u: %.1f ms, m %llu\n",
               double(t_end - t_first) * ms_per_tick, double(t_last_start - t_first) * ms_per_tick,
               busy * ms_per_tick * 1e-3, busy / double(std::max(1ull, t_end - t_first)), g_longest,
               double(longest) * ms_per_tick, (unsigned long long)(h_off[g_longest + 1] - h_off[g_longest]));
  std::fprintf(stderr, "[batch] phase ms summed per bucket: build core count1 peel1 count2 peel2 | "
                       "mean levels/rounds pass1 pass2 | max graph ms\n");
  for (uint32_t b = 0; b < 41; ++b) {
    const Acc& a = acc[b];
    if (!a.graphs) continue;
    std::fprintf(stderr,
                 "[batch] 2^%-2u %5u graphs: %9.1f %9.1f %9.1f %9.1f %9.1f %9.1f | %6.1f/%-7.1f %6.1f/%-7.1f | %8.1f\n",
                 b, a.graphs, a.ms[0], a.ms[1], a.ms[2], a.ms[3], a.ms[4], a.ms[5], a.lv[0] / a.graphs,
                 a.rd[0] / a.graphs, a.lv[1] / a.graphs, a.rd[1] / a.graphs, a.max_ms);
  }
}

// For every size bucket, up to `per_bucket` graphs of the batch are solved
// once per candidate lane count; the candidate with the smallest summed
// per-graph time wins the bucket.
void BatchRunner::autotune(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                           uint32_t per_bucket, uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status) {
  KT_HIP(hipSetDevice(device_));
  std::vector<std::vector<uint32_t>> by_bucket(41);
  for (uint32_t g = 0; g < count; ++g) {
    auto& v = by_bucket[bucket_of(h_off[g + 1] - h_off[g])];
    if (v.size() < per_bucket) v.push_back(g);
  }
  std::vector<uint32_t> sample;
  for (auto& v : by_bucket) sample.insert(sample.end(), v.begin(), v.end());
  if (sample.empty()) return;
  std::stable_sort(sample.begin(), sample.end(), [&](uint32_t a, uint32_t b) {
    return h_off[a + 1] - h_off[a] > h_off[b + 1] - h_off[b];
  });
  const uint8_t cand[] = {4, 8, 16, 32, 64};
  std::vector<std::vector<double>> cost(41, std::vector<double>(5, 0.0));
  ticks_.ensure(sizeof(unsigned long long) * std::max<size_t>(count, 1));
  std::vector<unsigned long long> ticks(count);
  for (int c = 0; c < 5; ++c) {
    std::vector<uint8_t> lanes(count, cand[c]);
    int grid = 0;
    BatchArgs a = plan(d_pairs, h_off, h_ns, count, sample, lanes, &grid);
    a.kmax = d_kmax;
    a.count = reinterpret_cast<unsigned long long*>(d_count);
    a.status = d_status;
    a.keep = nullptr;
    a.ticks = ticks_.as<unsigned long long>();
    batch_launch(a, grid, stream_);
    KT_HIP(hipMemcpyAsync(ticks.data(), ticks_.p, sizeof(unsigned long long) * count, hipMemcpyDeviceToHost,
                          stream_));
    KT_HIP(hipStreamSynchronize(stream_));
    for (uint32_t g : sample) cost[bucket_of(h_off[g + 1] - h_off[g])][c] += double(ticks[g]);
  }
  for (uint32_t b = 0; b < 41; ++b) {
    if (by_bucket[b].empty()) continue;
    int best = 0;
    for (int c = 1; c < 5; ++c)
      if (cost[b][c] < cost[b][best]) best = c;
    table_[b] = cand[best];
  }
  drop_graph();
}

void BatchRunner::table(uint8_t* out41) const { std::memcpy(out41, table_, sizeof(table_)); }

}  // namespace kt
