// Host side of the batched solve: validation, the unions of the valid graphs
// (one per part of the batch), their truss decompositions, per-graph answers.
#include "kt_batch_host.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <exception>
#include <thread>
#include <vector>

#include "kt_batch_union.h"
#include "kt_kernels.h"

namespace kt {

BatchRunner::BatchRunner(Solver* s, uint32_t max_parts) : solver_(s) {
  if (const char* v = std::getenv("KT_BATCH_PARTS")) parts_wanted_ = uint32_t(std::max(1, std::min(16, std::atoi(v))));
  if (max_parts) parts_wanted_ = max_parts;
}

BatchRunner::~BatchRunner() {
  (void)hipSetDevice(solver_->device());
  (void)hipStreamSynchronize(solver_->stream());
  release_buffers();
}

uint64_t BatchRunner::device_bytes() const {
  uint64_t b = 0;
  for (const DevBuf* d : {&poff_, &ns_, &status_, &gk_, &gcnt_}) b += d->bytes;
  for (const Part& p : parts_) {
    for (const DevBuf* d : {&p.poff, &p.vstart, &p.upoff, &p.uni}) b += d->bytes;
    if (p.own) b += p.own->device_bytes();
  }
  return b;
}

void BatchRunner::release_buffers() {
  for (DevBuf* d : {&poff_, &ns_, &status_, &gk_, &gcnt_}) d->release();
  for (Part& p : parts_) {
    if (p.own) (void)hipStreamSynchronize(p.own->stream());
    for (DevBuf* d : {&p.poff, &p.vstart, &p.upoff, &p.uni}) d->release();
  }
  parts_.clear();  // their solvers free their own buffers
}

// Graphs [g0, g1) of the batch as one union on solver sv (and its stream):
// union numbering of its valid graphs, the union, its truss decomposition
// and the per-graph answers into gk_ / gcnt_ (and keep) at [g0, g1). The
// part's pairs, offsets and results are slices of the batch's, so parts
// write disjoint slots.
void BatchRunner::solve_part(Part& part, Solver* sv, const uint32_t* d_pairs, const uint64_t* h_off,
                             const uint32_t* h_ns, const std::vector<uint32_t>& st, uint32_t g0, uint32_t g1,
                             uint8_t* d_keep) {
  KT_HIP(hipSetDevice(sv->device()));
  const hipStream_t s = sv->stream();
  const uint32_t cnt = g1 - g0;
  const uint64_t base = h_off[g0];
  std::vector<uint64_t> poff(size_t(cnt) + 1), upoff(cnt);
  std::vector<uint32_t> vstart(size_t(cnt) + 1);
  uint64_t nv = 0, mu = 0;
  for (uint32_t i = 0; i < cnt; ++i) {
    const uint32_t g = g0 + i;
    poff[i] = h_off[g] - base;
    vstart[i] = uint32_t(nv);
    upoff[i] = mu;
    if (st[g]) continue;
    nv += h_ns[g];
    mu += h_off[g + 1] - h_off[g];
    if (nv >= (uint64_t(1) << 32)) fail("batch: the union of the graphs has 2^32 or more vertices");
  }
  poff[cnt] = h_off[g1] - base;
  vstart[cnt] = uint32_t(nv);
  if (!mu) return;
  part.poff.ensure(sizeof(uint64_t) * (size_t(cnt) + 1));
  part.vstart.ensure(sizeof(uint32_t) * (size_t(cnt) + 1));
  part.upoff.ensure(sizeof(uint64_t) * cnt);
  part.uni.ensure(sizeof(uint32_t) * 2 * mu);
  KT_HIP(hipMemcpyAsync(part.poff.p, poff.data(), sizeof(uint64_t) * (size_t(cnt) + 1), hipMemcpyHostToDevice, s));
  KT_HIP(hipMemcpyAsync(part.vstart.p, vstart.data(), sizeof(uint32_t) * (size_t(cnt) + 1), hipMemcpyHostToDevice,
                        s));
  KT_HIP(hipMemcpyAsync(part.upoff.p, upoff.data(), sizeof(uint64_t) * cnt, hipMemcpyHostToDevice, s));
  const uint32_t* pairs = d_pairs + 2 * base;
  batch_union(pairs, part.poff.as<uint64_t>(), cnt, poff[cnt], part.vstart.as<uint32_t>(),
              status_.as<uint32_t>() + g0, part.upoff.as<uint64_t>(), part.uni.as<uint32_t>(), s);
  CoreView cv;
  const uint32_t kall = sv->decompose_pairs_device(part.uni.as<uint32_t>(), mu, uint32_t(nv), &cv);
  batch_truss(cv, kall, part.vstart.as<uint32_t>(), cnt, gk_.as<uint32_t>() + g0,
              reinterpret_cast<unsigned long long*>(gcnt_.p) + g0, pairs, part.poff.as<uint64_t>(),
              d_keep ? d_keep + base : nullptr, s);
  KT_HIP(hipStreamSynchronize(s));
}

void BatchRunner::solve(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                        uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep) {
  solve_impl(d_pairs, h_off, h_ns, count, d_kmax, d_count, d_status, d_keep, 1);
}

BatchTuning BatchRunner::autotune(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                                  uint32_t every, uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status) {
  if (every == 0) fail("batch autotune: sample_every must be at least 1");
  const hipStream_t s = solver_->stream();
  const Solver::Tuning entry = solver_->tuning();
  BatchTuning out;
  std::vector<uint32_t> k0(count), s0(count), k1(count), s1(count);
  std::vector<uint64_t> c0(count), c1(count);
  auto trial = [&](const Solver::Tuning& cand) {
    solver_->set_tuning(cand);
    const auto t0 = std::chrono::steady_clock::now();
    solve_impl(d_pairs, h_off, h_ns, count, d_kmax, d_count, d_status, nullptr, every);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (count) {
      KT_HIP(hipMemcpyAsync(k1.data(), d_kmax, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s));
      KT_HIP(hipMemcpyAsync(c1.data(), d_count, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, s));
      KT_HIP(hipMemcpyAsync(s1.data(), d_status, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s));
      KT_HIP(hipStreamSynchronize(s));
    }
    if (out.trials == 0) {
      k0.swap(k1), c0.swap(c1), s0.swap(s1);
    } else if (k1 != k0 || c1 != c0 || s1 != s0) {
      fail("batch autotune: two settings gave a graph different answers");
    }
    ++out.trials;
    return ms;
  };
  try {
    Solver::Tuning best = entry;
    trial(best);  // buffers grow to the batch
    double best_ms = trial(best);
    out.ms_default = best_ms;
    auto consider = [&](const Solver::Tuning& cand) {
      const double ms = trial(cand);
      if (ms < 0.97 * best_ms) best_ms = ms, best = cand;
    };
    const int widths[] = {8, 16, 32, 64};
    const Solver::Tuning base1 = best;
    for (int l : widths)
      if (l != base1.peel_lanes) {
        Solver::Tuning c = base1;
        c.peel_lanes = l;
        consider(c);
      }
    const Solver::Tuning base2 = best;
    for (int l : widths)
      if (l != base2.count_lanes) {
        Solver::Tuning c = base2;
        c.count_lanes = l;
        consider(c);
      }
    Solver::Tuning c3 = best;
    c3.edge_count = !c3.edge_count;
    consider(c3);
    const Solver::Tuning base4 = best;
    for (double f : {0.5, 2.0, 4.0}) {
      Solver::Tuning c = base4;
      c.drop_ratio = base4.drop_ratio * f;
      consider(c);
    }
    solver_->set_tuning(best);
    out.best = best;
    out.ms_best = best_ms;
    solve_impl(d_pairs, h_off, h_ns, count, d_kmax, d_count, d_status, nullptr, 1);
  } catch (...) {
    solver_->set_tuning(entry);
    throw;
  }
  return out;
}

void BatchRunner::solve_impl(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                             uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep,
                             uint32_t every) {
  KT_HIP(hipSetDevice(solver_->device()));
  if (!count) return;
  const hipStream_t s = solver_->stream();
  for (uint32_t g = 0; g < count; ++g)
    if (h_off[g + 1] < h_off[g]) fail("batch: offsets must be non-decreasing");
  const uint64_t total = h_off[count];
  if (total && !d_pairs) fail("batch: NULL pairs");
  poff_.ensure(sizeof(uint64_t) * (size_t(count) + 1));
  ns_.ensure(sizeof(uint32_t) * count);
  status_.ensure(sizeof(uint32_t) * count);
  KT_HIP(hipMemcpyAsync(poff_.p, h_off, sizeof(uint64_t) * (size_t(count) + 1), hipMemcpyHostToDevice, s));
  KT_HIP(hipMemcpyAsync(ns_.p, h_ns, sizeof(uint32_t) * count, hipMemcpyHostToDevice, s));
  fill_u32(status_.as<uint32_t>(), count, 0, s);
  batch_validate(d_pairs, poff_.as<uint64_t>(), ns_.as<uint32_t>(), count, total, status_.as<uint32_t>(), s);
  std::vector<uint32_t> st(count);
  KT_HIP(hipMemcpyAsync(st.data(), status_.p, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s));
  KT_HIP(hipStreamSynchronize(s));
  if (every > 1) {  // a sample: the other graphs take no part
    for (uint32_t g = 0; g < count; ++g)
      if (g % every) st[g] = 2;
    KT_HIP(hipMemcpyAsync(status_.p, st.data(), sizeof(uint32_t) * count, hipMemcpyHostToDevice, s));
  }

  gk_.ensure(sizeof(uint32_t) * count);
  gcnt_.ensure(sizeof(unsigned long long) * count);
  fill_u32(gk_.as<uint32_t>(), count, 0, s);
  fill_u64(gcnt_.as<uint64_t>(), count, 0, s);
  if (d_keep) fill_u8(d_keep, total, 0, s);
  // parts: contiguous graph ranges of about equal valid edges, each one
  // union decomposed on its own solver and stream, all at once (an invalid
  // graph takes no vertex and no edge)
  uint64_t valid = 0;
  for (uint32_t g = 0; g < count; ++g)
    if (!st[g]) valid += h_off[g + 1] - h_off[g];
  const uint32_t np = valid < (uint64_t(1) << 16) ? 1u : std::max(1u, std::min(parts_wanted_, count));
  std::vector<uint32_t> cut(size_t(np) + 1, count);
  cut[0] = 0;
  {
    uint64_t acc = 0;
    uint32_t p = 1;
    for (uint32_t g = 0; g < count && p < np; ++g) {
      if (!st[g]) acc += h_off[g + 1] - h_off[g];
      if (acc * np >= valid * p) cut[p++] = g + 1;
    }
  }
  while (parts_.size() < np) parts_.emplace_back();
  const Solver::Tuning tune = solver_->tuning();
  for (uint32_t p = 1; p < np; ++p) {
    if (!parts_[p].own) parts_[p].own.reset(new Solver(solver_->device()));
    parts_[p].own->set_tuning(tune);
  }
  KT_HIP(hipStreamSynchronize(s));  // zeroed results and the status before any part reads them
  std::vector<std::exception_ptr> errs(np);
  std::vector<std::thread> threads;
  for (uint32_t p = 1; p < np; ++p)
    threads.emplace_back([&, p] {
      try {
        solve_part(parts_[p], parts_[p].own.get(), d_pairs, h_off, h_ns, st, cut[p], cut[p + 1], d_keep);
      } catch (...) {
        errs[p] = std::current_exception();
      }
    });
  try {
    solve_part(parts_[0], solver_, d_pairs, h_off, h_ns, st, cut[0], cut[1], d_keep);
  } catch (...) {
    errs[0] = std::current_exception();
  }
  for (std::thread& t : threads) t.join();
  for (const std::exception_ptr& e : errs)
    if (e) std::rethrow_exception(e);
  batch_results(gk_.as<uint32_t>(), reinterpret_cast<const unsigned long long*>(gcnt_.p), status_.as<uint32_t>(),
                count, d_kmax, reinterpret_cast<unsigned long long*>(d_count), d_status, s);
  KT_HIP(hipStreamSynchronize(s));
}

}  // namespace kt
