// Host orchestration of the batched solve (one per context): the valid graphs
// of a batch become one disjoint union whose truss decomposition the
// context's solver computes; per-graph answers are read off it.
#pragma once
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
  explicit BatchRunner(Solver* s);
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
  // device bytes of its own buffers (the solver's are counted by the solver)
  uint64_t device_bytes() const;
  void release_buffers();

 private:
  // every > 1: graphs g with g % every != 0 take no part (status 2)
  void solve_impl(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                  uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep, uint32_t every);
  Solver* solver_;
  DevBuf poff_, ns_, status_, vstart_, upoff_, union_, gk_, gcnt_;
};

}  // namespace kt
