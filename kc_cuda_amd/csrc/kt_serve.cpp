// Serving loop: admission by HBM footprint, one solver context (stream +
// host thread) per slot, results handed back by ticket.
#include "kt_serve.h"

#include <algorithm>
#include <cstdlib>

namespace kt {

namespace {
double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}
}  // namespace

// Per input edge: the pairs' device copy, the CSR and its build and sort
// buffers, the vertex-peel nest's subgraphs, and for every core edge its
// oriented and undirected slots with edge ids, support, state, deletion
// level, queues, backups and both compaction sets (160 B in all; measured at
// most 96 B on serving and batch mixes and RMAT scale 22); per vertex
// offsets, degrees, levels and maps (64 B); a dense core's bit rows (at most
// 72 B per edge, since a row is no longer than two sorted rows) and its
// edge-id matrix (at most kServeEidmBytes); 16 MiB of fixed buffers.
uint64_t Server::footprint(uint64_t m, uint32_t n) {
  return 160ull * m + 64ull * uint64_t(n) + std::min<uint64_t>(72ull * m, kServeEidmBytes) +
         std::min<uint64_t>(512ull * m, kServeEidmBytes) + (16ull << 20);
}

Server::Server(int device, int slots, uint64_t budget) : device_(device), budget_(budget) {
  if (slots < 1 || slots > 64) fail("serve: slots must be in [1, 64]");
  if (const char* v = std::getenv("KT_SERVE_GROUP_EDGES")) group_query_edges_ = uint64_t(std::atoll(v));
  if (const char* v = std::getenv("KT_SERVE_GROUP_TOTAL")) group_edges_ = uint64_t(std::atoll(v));
  KT_HIP(hipSetDevice(device_));
  slots_.resize(size_t(slots));
  for (auto& s : slots_) {
    s.solver = std::make_unique<Solver>(device_);
    s.solver->set_eidm_budget(kServeEidmBytes);
    s.batch = std::make_unique<BatchRunner>(s.solver.get(), 1);
    s.held = s.charge = slot_bytes(s);
    ctr_.bytes_in_use += s.charge;
  }
  ctr_.peak_bytes_in_use = ctr_.bytes_in_use;
  for (uint32_t i = 0; i < uint32_t(slots); ++i) slots_[i].thread = std::thread([this, i] { run_slot(i); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_admit_.notify_all();
  cv_done_.notify_all();
  for (auto& s : slots_)
    if (s.thread.joinable()) s.thread.join();
  for (auto& s : slots_) {
    s.batch.reset();
    s.staging.release();
    s.gres.release();
  }
}

uint64_t Server::submit(const uint32_t* pairs, uint64_t m, uint32_t n, bool on_device) {
  if (m && !pairs) fail("serve: NULL pairs");
  Query q{0, pairs, m, n, on_device, footprint(m, n), Clock::now()};
  std::unique_lock<std::mutex> lk(mu_);
  q.id = next_id_++;
  ctr_.submitted += 1;
  open_.insert(q.id);
  if (q.bytes + uint64_t(slots_.size()) * kServeSlotFloor > budget_) {  // can never be admitted
    ServeOutcome out;
    out.status = 1;
    out.footprint = q.bytes;
    out.error = "footprint " + std::to_string(q.bytes) + " B exceeds the serving budget of " +
                std::to_string(budget_) + " B";
    ctr_.rejected += 1;
    done_.emplace(q.id, std::move(out));
    lk.unlock();
    cv_done_.notify_all();
    return q.id;
  }
  queue_.push_back(q);
  lk.unlock();
  cv_admit_.notify_all();
  return q.id;
}

ServeOutcome Server::wait(uint64_t ticket) {
  std::unique_lock<std::mutex> lk(mu_);
  if (!open_.count(ticket)) fail("serve: unknown ticket, or already waited for");
  cv_done_.wait(lk, [&] { return done_.count(ticket) != 0 || stop_; });
  if (!done_.count(ticket)) fail("serve: server stopped");
  ServeOutcome out = std::move(done_[ticket]);
  done_.erase(ticket);
  open_.erase(ticket);
  return out;
}

ServeCounters Server::counters() {
  std::lock_guard<std::mutex> lk(mu_);
  return ctr_;
}

void Server::finish(uint32_t s, std::vector<Query>& group, std::vector<ServeOutcome>& outs) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    Slot& slot = slots_[s];
    for (size_t i = 0; i < group.size(); ++i) {
      ServeOutcome& out = outs[i];
      ctr_.in_flight -= 1;
      if (out.status == 0) {
        ctr_.completed += 1;
        if (out.slot_bytes > out.charged) ctr_.over_estimate += 1;
        if (out.group > 1) ctr_.grouped += 1;
      } else {
        ctr_.failed += 1;
      }
      done_.emplace(group[i].id, std::move(out));
    }
    // from now on the slot is charged what it holds
    ctr_.bytes_in_use = ctr_.bytes_in_use - slot.charge + slot.held;
    slot.charge = slot.held;
    ctr_.peak_bytes_in_use = std::max(ctr_.peak_bytes_in_use, ctr_.bytes_in_use);
    busy_ -= 1;
  }
  cv_admit_.notify_all();  // memory and a slot freed up
  cv_done_.notify_all();
}

// Continuous batching (lock held, group = {head}, head already charged):
// while more queries wait than the idle slots can take, the small ones right
// behind the head join its group if their footprints fit in the budget too.
// A big query ends the group; queries keep their order.
void Server::take_group(Slot& slot, std::vector<Query>& group) {
  if (group[0].m > group_query_edges_) return;
  const size_t idle = slots_.size() - busy_;
  uint64_t edges = group[0].m, verts = group[0].n;
  while (group.size() < kGroupMax && queue_.size() > idle) {
    const Query& q = queue_.front();
    if (q.m > group_query_edges_ || edges + q.m > group_edges_) break;
    if (verts + q.n >= (uint64_t(1) << 32)) break;  // union ids are 32-bit
    if (ctr_.bytes_in_use + q.bytes > budget_) break;
    edges += q.m;
    verts += q.n;
    ctr_.bytes_in_use += q.bytes;
    slot.charge += q.bytes;
    group.push_back(q);
    queue_.pop_front();
  }
}

void Server::solve_alone(Slot& slot, const Query& q, ServeOutcome& out) {
  try {
    Solver& sv = *slot.solver;
    const uint32_t* d = q.pairs;
    if (!q.on_device) {
      slot.staging.ensure(sizeof(uint32_t) * 2 * std::max<uint64_t>(q.m, 1));
      if (q.m)
        KT_HIP(hipMemcpyAsync(slot.staging.p, q.pairs, sizeof(uint32_t) * 2 * q.m, hipMemcpyHostToDevice,
                              sv.stream()));
      d = slot.staging.as<uint32_t>();
    }
    sv.solve_pairs_device(d, q.m, q.n, &out.res);
    KT_HIP(hipStreamSynchronize(sv.stream()));
  } catch (const Error& e) {
    out.status = 2;
    out.error = e.what;
  } catch (const std::exception& e) {
    out.status = 2;
    out.error = e.what();
  }
}

// The group's pairs are gathered into the staging buffer and solved by the
// batched solve (one truss decomposition of their disjoint union). A query the
// batched solve flags (pairs not canonical and sorted) is marked for
// solve_alone, which gives it the stand-alone answer or error; so is every
// query when the group solve itself fails.
void Server::solve_group(Slot& slot, const std::vector<Query>& group, std::vector<ServeOutcome>& outs,
                         std::vector<char>& alone) {
  const uint32_t count = uint32_t(group.size());
  try {
    Solver& sv = *slot.solver;
    const hipStream_t st = sv.stream();
    std::vector<uint64_t> off(size_t(count) + 1, 0);
    std::vector<uint32_t> ns(count);
    for (uint32_t i = 0; i < count; ++i) {
      off[i + 1] = off[i] + group[i].m;
      ns[i] = group[i].n;
    }
    slot.staging.ensure(sizeof(uint32_t) * 2 * std::max<uint64_t>(off[count], 1));
    for (uint32_t i = 0; i < count; ++i)
      if (group[i].m)
        KT_HIP(hipMemcpyAsync(slot.staging.as<uint32_t>() + 2 * off[i], group[i].pairs,
                              sizeof(uint32_t) * 2 * group[i].m,
                              group[i].on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    slot.gres.ensure((sizeof(uint64_t) + 2 * sizeof(uint32_t)) * count);
    uint64_t* d_cnt = slot.gres.as<uint64_t>();
    uint32_t* d_kmax = reinterpret_cast<uint32_t*>(d_cnt + count);
    uint32_t* d_status = d_kmax + count;
    slot.batch->solve(slot.staging.as<uint32_t>(), off.data(), ns.data(), count, d_kmax, d_cnt, d_status, nullptr);
    std::vector<uint64_t> cnt(count);
    std::vector<uint32_t> kmax(count), status(count);
    KT_HIP(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, st));
    KT_HIP(hipMemcpyAsync(kmax.data(), d_kmax, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, st));
    KT_HIP(hipMemcpyAsync(status.data(), d_status, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, st));
    KT_HIP(hipStreamSynchronize(st));
    const uint32_t in_union = uint32_t(std::count(status.begin(), status.end(), 0u));
    for (uint32_t i = 0; i < count; ++i) {
      if (status[i]) {
        alone[i] = 1;
        continue;
      }
      kt_result& r = outs[i].res;
      r.n = ns[i];
      r.m = group[i].m;
      r.kmax = kmax[i];
      r.truss_edges = cnt[i];
      outs[i].group = in_union;
    }
  } catch (const Error&) {
    std::fill(alone.begin(), alone.end(), 1);
  } catch (const std::exception&) {
    std::fill(alone.begin(), alone.end(), 1);
  }
}

// One slot: take the head of the queue once its charge fits in the budget
// (first come, first served: a query that does not fit holds back the ones
// behind it), with the small queries that join its group, solve on this
// slot's stream, report. When nothing runs and the head still does not fit,
// idle slots free their buffers one by one.
void Server::run_slot(uint32_t s) {
  Slot& slot = slots_[s];
  (void)hipSetDevice(device_);
  std::vector<Query> group;
  std::vector<ServeOutcome> outs;
  std::vector<char> alone;
  for (;;) {
    bool trim = false;
    group.clear();
    {
      std::unique_lock<std::mutex> lk(mu_);
      auto fits = [&] {
        const Query& h = queue_.front();
        return ctr_.bytes_in_use + h.bytes <= budget_;  // the slot may grow by the footprint
      };
      cv_admit_.wait(lk, [&] {
        if (stop_) return true;
        if (queue_.empty()) return false;
        if (fits()) return true;
        return ctr_.in_flight == 0 && slot.held > kServeSlotFloor;  // nothing will free memory: trim
      });
      if (stop_) return;
      if (!fits()) {
        trim = true;
      } else {
        group.push_back(queue_.front());
        queue_.pop_front();
        busy_ += 1;
        // its buffers hold slot.held now and grow by at most the queries' footprints
        const uint64_t charge = slot.held + group[0].bytes;
        ctr_.bytes_in_use = ctr_.bytes_in_use - slot.charge + charge;
        slot.charge = charge;
        take_group(slot, group);
        ctr_.in_flight += group.size();
        ctr_.peak_in_flight = std::max(ctr_.peak_in_flight, ctr_.in_flight);
        ctr_.peak_bytes_in_use = std::max(ctr_.peak_bytes_in_use, ctr_.bytes_in_use);
      }
    }
    if (trim) {
      try {
        slot.solver->release_buffers();
        slot.batch->release_buffers();
        slot.staging.release();
        slot.gres.release();
      } catch (const Error&) {
        // a buffer that could not be freed stays counted in held
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        ctr_.trims += 1;
        slot.held = slot_bytes(slot);
        ctr_.bytes_in_use = ctr_.bytes_in_use - slot.charge + slot.held;
        slot.charge = slot.held;
      }
      cv_admit_.notify_all();
      continue;
    }
    cv_admit_.notify_all();  // the next query may fit beside these
    const size_t count = group.size();
    outs.assign(count, ServeOutcome{});
    alone.assign(count, count == 1 ? 1 : 0);
    const auto t_admit = Clock::now();
    for (size_t i = 0; i < count; ++i) {
      outs[i].slot = s;
      outs[i].footprint = group[i].bytes;
      outs[i].charged = slot.charge;
      outs[i].queued_ms = ms_between(group[i].t_submit, t_admit);
    }
    if (count > 1) solve_group(slot, group, outs, alone);
    for (size_t i = 0; i < count; ++i)
      if (alone[i]) solve_alone(slot, group[i], outs[i]);
    const auto t_done = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot.held = slot_bytes(slot);
    }
    for (size_t i = 0; i < count; ++i) {
      outs[i].slot_bytes = slot.held;
      outs[i].run_ms = ms_between(t_admit, t_done);
      outs[i].total_ms = ms_between(group[i].t_submit, t_done);
    }
    finish(s, group, outs);
  }
}

}  // namespace kt
