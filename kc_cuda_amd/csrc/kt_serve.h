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

#include "kt_batch_host.h"
#include "kt_solver.h"

namespace kt {

// Edge-id matrix cap of a serving slot's dense cores (bytes).
constexpr uint64_t kServeEidmBytes = uint64_t(256) << 20;
// Bytes an idle slot may keep after freeing its buffers (control block).
constexpr uint64_t kServeSlotFloor = uint64_t(1) << 20;
// Continuous batching: while more queries wait than the idle slots can take,
// a slot that takes a query of at most kGroupQueryEdges edges also takes the
// small queries queued right behind it (up to kGroupMax queries and
// kGroupEdges edges in all, each charged its footprint within the budget)
// and solves them as one disjoint union (one truss decomposition). The union
// walks every truss level of every graph, where a stand-alone solve skips to
// the top ones, so it pays only where launch latency dominates a solve: on
// 8 slots it serves 512 queries of median 20k edges 23% faster and 160 of
// median 200k edges at the same rate; a per-query cap of 2^18 edges is 90%
// faster on the first and 30% slower on the second. KT_SERVE_GROUP_EDGES
// (0: no grouping) and KT_SERVE_GROUP_TOTAL override the caps.
constexpr uint64_t kGroupQueryEdges = uint64_t(1) << 16;
constexpr uint64_t kGroupEdges = uint64_t(4) << 20;
constexpr size_t kGroupMax = 64;

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
  uint32_t group = 1;         // queries solved together in one union (1: alone)
  std::string error;
};

struct ServeCounters {
  uint64_t submitted = 0, completed = 0, rejected = 0, failed = 0;
  uint64_t in_flight = 0, peak_in_flight = 0;
  uint64_t bytes_in_use = 0, peak_bytes_in_use = 0;  // charged to the slots, idle ones included
  uint64_t over_estimate = 0;  // solves whose slot grew past the bytes charged for them
  uint64_t trims = 0;          // idle slots that freed their buffers to let a query in
  uint64_t grouped = 0;        // queries solved in a union with others
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
    std::unique_ptr<BatchRunner> batch;  // a group's union solve, on the same solver
    DevBuf staging;      // device copy of a host query's pairs (a group's, concatenated)
    DevBuf gres;         // a group's per-query kmax, |T| and status
    uint64_t held = 0;   // device bytes its buffers hold (set by its own thread)
    uint64_t charge = 0; // bytes counted in bytes_in_use for it
    std::thread thread;
  };
  uint64_t slot_bytes(Slot& s) const {
    return s.solver->device_bytes() + s.batch->device_bytes() + s.staging.bytes + s.gres.bytes;
  }
  void run_slot(uint32_t s);
  // takes the queries that may join a group headed by q (lock held)
  void take_group(Slot& slot, std::vector<Query>& group);
  void solve_alone(Slot& slot, const Query& q, ServeOutcome& out);
  // solves the group as one union; a query whose pairs are not canonical and
  // sorted is left for solve_alone (its flag in alone is set)
  void solve_group(Slot& slot, const std::vector<Query>& group, std::vector<ServeOutcome>& outs,
                   std::vector<char>& alone);
  void finish(uint32_t s, std::vector<Query>& group, std::vector<ServeOutcome>& outs);

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
  uint32_t busy_ = 0;  // slots solving
  uint64_t group_query_edges_ = kGroupQueryEdges;  // 0: no grouping
  uint64_t group_edges_ = kGroupEdges;
  bool stop_ = false;
};

}  // namespace kt
