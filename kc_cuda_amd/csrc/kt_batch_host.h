// Host orchestration of the batched solve (one per context): the valid graphs
// of a batch become disjoint unions (one per part of the batch, decomposed
// concurrently) whose truss decompositions give the per-graph answers.
#pragma once
#include <memory>
#include <vector>

#include "kt_solver.h"

namespace kt {

struct BatchTuning {
  Solver::Tuning best;
  double ms_default = 0;  // sampled batch with the solver's settings on entry
  double ms_best = 0;     // with best
  uint32_t trials = 0;    // timed batch solves
};

class BatchRunner {
 public:
  // max_parts: concurrent parts a batch may be split into (0: KT_BATCH_PARTS
  // or 4; serving slots pass 1, their groups sharing the slot's solver only)
  explicit BatchRunner(Solver* s, uint32_t max_parts = 0);
  ~BatchRunner();
  // d_pairs: all graphs' canonical pairs (sorted within each graph) in HBM;
  // h_off / h_ns: host offsets (count + 1) and vertex counts. Results go to
  // device arrays; d_keep (nullable) gets 1 for every pair in its graph's
  // kmax-truss.
  void solve(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
             uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep);
  // Autotuner: times the batched solve of every `every`-th graph under
  // candidate solver settings, one setting at a time (peel lanes, count
  // lanes, count kernel, drop ratio; a candidate is kept when it is 3%
  // faster), checks that every trial gives every graph the same answer,
  // leaves the fastest settings on the solver and solves the whole batch
  // into the outputs.
  BatchTuning autotune(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                       uint32_t every, uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status);
  // device bytes of its own buffers and of its extra parts' solvers (the
  // context solver's are counted by the solver)
  uint64_t device_bytes() const;
  void release_buffers();

 private:
  // A part of the batch: a contiguous range of graphs whose union one
  // solver (the context's for part 0, an own one with its own stream for
  // the others) decomposes; the parts run concurrently.
  struct Part {
    std::unique_ptr<Solver> own;
    DevBuf poff, vstart, upoff, uni;
  };
  // every > 1: graphs g with g % every != 0 take no part (status 2)
  void solve_impl(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                  uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep, uint32_t every);
  void solve_part(Part& part, Solver* sv, const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns,
                  const std::vector<uint32_t>& st, uint32_t g0, uint32_t g1, uint8_t* d_keep);
  Solver* solver_;
  DevBuf poff_, ns_, status_, gk_, gcnt_;
  std::vector<Part> parts_;
  uint32_t parts_wanted_ = 4;  // concurrent parts of a batch (1: one union)
};

}  // namespace kt
