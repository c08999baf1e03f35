// Host orchestration of the batched solve (one per context).
#pragma once
#include <vector>

#include "kt_batch.h"
#include "kt_solver.h"

namespace kt {

class BatchRunner {
 public:
  explicit BatchRunner(Solver* s);
  ~BatchRunner();
  // d_pairs: all graphs' canonical pairs (sorted within each graph) in HBM;
  // h_off / h_ns: host offsets (count + 1) and vertex counts. Results go to
  // device arrays. use_graph: launch through a captured hipGraph, replayed
  // while the arguments stay the same.
  void solve(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
             uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep, bool use_graph);
  // Fills the lanes-per-edge table from timed runs on a sample of the batch
  // (results of the sample runs land in the given arrays).
  void autotune(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                uint32_t per_bucket, uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status);
  void table(uint8_t* out41) const;

 private:
  uint8_t lanes_for(uint64_t m, uint32_t n) const;
  BatchArgs plan(const uint32_t* d_pairs, const uint64_t* h_off, const uint32_t* h_ns, uint32_t count,
                 const std::vector<uint32_t>& order, const std::vector<uint8_t>& lanes, int* grid);
  void drop_graph();
  void report_phases(const uint64_t* h_off, uint32_t count);

  hipStream_t stream_;
  int device_;
  int cus_ = 0;
  DevBuf poff_, ns_, order_, lanes_, slab_, work_, ticks_, phases_;
  bool profile_ = false;  // KT_BATCH_PROFILE=1: per-phase times by size bucket on stderr
  uint8_t table_[41];  // lanes per edge by bucket floor(log2 m); 0 = by degree
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  BatchArgs cached_{};
  int cached_grid_ = 0;
};

}  // namespace kt
