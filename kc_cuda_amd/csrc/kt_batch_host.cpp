// Host side of the batched solve: validation, the union of the valid graphs,
// one truss decomposition, per-graph answers.
#include "kt_batch_host.h"

#include <algorithm>
#include <chrono>
#include <vector>

#include "kt_batch_union.h"
#include "kt_kernels.h"

namespace kt {

BatchRunner::BatchRunner(Solver* s) : solver_(s) {}

BatchRunner::~BatchRunner() {
  (void)hipSetDevice(solver_->device());
  (void)hipStreamSynchronize(solver_->stream());
  release_buffers();
}

uint64_t BatchRunner::device_bytes() const {
  uint64_t b = 0;
  for (const DevBuf* d : {&poff_, &ns_, &status_, &vstart_, &upoff_, &union_, &gk_, &gcnt_}) b += d->bytes;
  return b;
}

void BatchRunner::release_buffers() {
  for (DevBuf* d : {&poff_, &ns_, &status_, &vstart_, &upoff_, &union_, &gk_, &gcnt_}) d->release();
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

  // union numbering of the valid graphs (an invalid graph takes no vertex)
  std::vector<uint32_t> vstart(size_t(count) + 1);
  std::vector<uint64_t> upoff(count);
  uint64_t nv = 0, mu = 0;
  for (uint32_t g = 0; g < count; ++g) {
    vstart[g] = uint32_t(nv);
    upoff[g] = mu;
    if (st[g]) continue;
    nv += h_ns[g];
    mu += h_off[g + 1] - h_off[g];
    if (nv >= (uint64_t(1) << 32)) fail("batch: the union of the graphs has 2^32 or more vertices");
  }
  vstart[count] = uint32_t(nv);
  vstart_.ensure(sizeof(uint32_t) * (size_t(count) + 1));
  upoff_.ensure(sizeof(uint64_t) * count);
  gk_.ensure(sizeof(uint32_t) * count);
  gcnt_.ensure(sizeof(unsigned long long) * count);
  KT_HIP(hipMemcpyAsync(vstart_.p, vstart.data(), sizeof(uint32_t) * (size_t(count) + 1), hipMemcpyHostToDevice, s));
  KT_HIP(hipMemcpyAsync(upoff_.p, upoff.data(), sizeof(uint64_t) * count, hipMemcpyHostToDevice, s));
  fill_u32(gk_.as<uint32_t>(), count, 0, s);
  fill_u64(gcnt_.as<uint64_t>(), count, 0, s);
  if (d_keep) fill_u8(d_keep, total, 0, s);
  if (mu) {
    union_.ensure(sizeof(uint32_t) * 2 * mu);
    batch_union(d_pairs, poff_.as<uint64_t>(), count, total, vstart_.as<uint32_t>(), status_.as<uint32_t>(),
                upoff_.as<uint64_t>(), union_.as<uint32_t>(), s);
    CoreView cv;
    const uint32_t kall = solver_->decompose_pairs_device(union_.as<uint32_t>(), mu, uint32_t(nv), &cv);
    batch_truss(cv, kall, vstart_.as<uint32_t>(), count, gk_.as<uint32_t>(),
                reinterpret_cast<unsigned long long*>(gcnt_.p), d_pairs, poff_.as<uint64_t>(), d_keep, s);
  }
  batch_results(gk_.as<uint32_t>(), reinterpret_cast<const unsigned long long*>(gcnt_.p), status_.as<uint32_t>(),
                count, d_kmax, reinterpret_cast<unsigned long long*>(d_count), d_status, s);
  KT_HIP(hipStreamSynchronize(s));
}

}  // namespace kt
