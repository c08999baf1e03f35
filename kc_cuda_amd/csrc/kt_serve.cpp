// Serving loop: admission by HBM footprint, one solver context (stream +
// host thread) per slot, results handed back by ticket.
#include "kt_serve.h"

#include <algorithm>

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
  KT_HIP(hipSetDevice(device_));
  slots_.resize(size_t(slots));
  for (auto& s : slots_) {
    s.solver = std::make_unique<Solver>(device_);
    s.solver->set_eidm_budget(kServeEidmBytes);
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
  for (auto& s : slots_) s.staging.release();
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

void Server::finish(uint32_t s, uint64_t id, ServeOutcome&& out) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    Slot& slot = slots_[s];
    ctr_.in_flight -= 1;
    if (out.status == 0) {
      ctr_.completed += 1;
      if (out.slot_bytes > out.charged) ctr_.over_estimate += 1;
    } else {
      ctr_.failed += 1;
    }
    // from now on the slot is charged what it holds
    ctr_.bytes_in_use = ctr_.bytes_in_use - slot.charge + slot.held;
    slot.charge = slot.held;
    ctr_.peak_bytes_in_use = std::max(ctr_.peak_bytes_in_use, ctr_.bytes_in_use);
    done_.emplace(id, std::move(out));
  }
  cv_admit_.notify_all();  // memory freed up
  cv_done_.notify_all();
}

// One slot: take the head of the queue once its charge fits in the budget
// (first come, first served: a query that does not fit holds back the ones
// behind it), solve it on this slot's stream, report. When nothing runs and
// the head still does not fit, idle slots free their buffers one by one.
void Server::run_slot(uint32_t s) {
  Slot& slot = slots_[s];
  (void)hipSetDevice(device_);
  for (;;) {
    Query q;
    bool trim = false;
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
        q = queue_.front();
        queue_.pop_front();
        ctr_.in_flight += 1;
        ctr_.peak_in_flight = std::max(ctr_.peak_in_flight, ctr_.in_flight);
        // its buffers hold slot.held now and grow by at most the query's footprint
        const uint64_t charge = slot.held + q.bytes;
        ctr_.bytes_in_use = ctr_.bytes_in_use - slot.charge + charge;
        slot.charge = charge;
        ctr_.peak_bytes_in_use = std::max(ctr_.peak_bytes_in_use, ctr_.bytes_in_use);
      }
    }
    if (trim) {
      std::string err;
      try {
        slot.solver->release_buffers();
        slot.staging.release();
      } catch (const Error& e) {
        err = e.what;
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
    cv_admit_.notify_all();  // the next query may fit beside this one
    ServeOutcome out;
    out.slot = s;
    out.footprint = q.bytes;
    out.charged = slot.charge;
    const auto t_admit = Clock::now();
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
    const auto t_done = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot.held = slot_bytes(slot);
    }
    out.slot_bytes = slot.held;
    out.queued_ms = ms_between(q.t_submit, t_admit);
    out.run_ms = ms_between(t_admit, t_done);
    out.total_ms = ms_between(q.t_submit, t_done);
    finish(s, q.id, std::move(out));
  }
}

}  // namespace kt
