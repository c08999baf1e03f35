// Serving loop: k-max-truss queries admitted by HBM footprint onto a fixed
// set of slots, each slot a solver context with its own HIP stream and host
// thread, so that queries overlap on the device (continuous batching: a query
// takes the first slot and memory that free up). No reference counterpart;
// the reference answers one graph per process (src/kc_gpu.cpp:209-347).
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "kt_solver.h"

namespace kt {

// Edge-id matrix cap of a serving slot's dense cores (bytes).
constexpr uint64_t kServeEidmBytes = uint64_t(256) << 20;
// Bytes an idle slot may keep after freeing its buffers (control block).
constexpr uint64_t kServeSlotFloor = uint64_t(1) << 20;

struct ServeOutcome {
  int status = 0;  // 0 solved; 1 rejected (footprint above the budget); 2 solve failed
  kt_result res{};
  uint32_t slot = 0;
  uint64_t footprint = 0;     // footprint(m, n) of the query
  uint64_t charged = 0;       // bytes charged to its slot while it ran
  uint64_t slot_bytes = 0;    // device bytes its slot held after the solve
  double queued_ms = 0;       // submit -> admitted
  double run_ms = 0;          // admitted -> solved (upload included)
  double total_ms = 0;        // submit -> solved
  std::string error;
};

struct ServeCounters {
  uint64_t submitted = 0, completed = 0, rejected = 0, failed = 0;
  uint64_t in_flight = 0, peak_in_flight = 0;
  uint64_t bytes_in_use = 0, peak_bytes_in_use = 0;  // charged to the slots, idle ones included
  uint64_t over_estimate = 0;  // solves whose slot grew past the bytes charged for them
  uint64_t trims = 0;          // idle slots that freed their buffers to let a query in
};

class Server {
 public:
  // slots: solver contexts (and streams); budget: HBM bytes the slots may
  // hold in total. A slot is charged what its buffers hold while idle and
  // that plus the footprint of its query while it runs (a buffer grows only
  // to what the query needs).
  Server(int device, int slots, uint64_t budget);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Queues one graph (canonical pairs u < v < n). The pairs (host or device
  // memory) must stay valid until wait() has returned for the ticket.
  uint64_t submit(const uint32_t* pairs, uint64_t m, uint32_t n, bool on_device);
  // Blocks until the query is solved (or rejected) and hands its outcome over.
  ServeOutcome wait(uint64_t ticket);
  ServeCounters counters();
  int slots() const { return int(slots_.size()); }

  // Admission charge of a query of m edges on n vertices: an upper bound on
  // the device bytes one solver context holds for it.
  static uint64_t footprint(uint64_t m, uint32_t n);

 private:
  using Clock = std::chrono::steady_clock;
  struct Query {
    uint64_t id;
    const uint32_t* pairs;
    uint64_t m;
    uint32_t n;
    bool on_device;
    uint64_t bytes;
    Clock::time_point t_submit;
  };
  struct Slot {
    std::unique_ptr<Solver> solver;
    DevBuf staging;      // device copy of a host query's pairs
    uint64_t held = 0;   // device bytes its buffers hold (set by its own thread)
    uint64_t charge = 0; // bytes counted in bytes_in_use for it
    std::thread thread;
  };
  uint64_t slot_bytes(Slot& s) const { return s.solver->device_bytes() + s.staging.bytes; }
  void run_slot(uint32_t s);
  void finish(uint32_t s, uint64_t id, ServeOutcome&& out);

  int device_;
  uint64_t budget_;
  std::vector<Slot> slots_;
  std::mutex mu_;
  std::condition_variable cv_admit_, cv_done_;
  std::deque<Query> queue_;
  std::unordered_map<uint64_t, ServeOutcome> done_;
  std::unordered_set<uint64_t> open_;  // submitted, not yet waited for
  ServeCounters ctr_;
  uint64_t next_id_ = 1;
  bool stop_ = false;
};

}  // namespace kt
