// Host side of the batched solve: plan (processing order, lanes per graph),
// scratch slabs, autotuner and hipGraph capture of the persistent launch.
#include "kt_batch_host.h"

#include <algorithm>
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
}

BatchRunner::~BatchRunner() {
  (void)hipSetDevice(device_);
  (void)hipStreamSynchronize(stream_);
  drop_graph();
  DevBuf* all[] = {&poff_, &ns_, &order_, &lanes_, &slab_, &work_, &ticks_};
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
  if (!use_graph) {
    batch_launch(a, grid, stream_);
    KT_HIP(hipStreamSynchronize(stream_));
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
