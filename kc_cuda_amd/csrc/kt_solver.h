// Host orchestration of the device-resident k-max-truss solve.
#pragma once
#include <vector>

#include "kt_common.h"
#include "../../include/ktruss.h"

namespace kt {

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b);
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Core subgraph in rank ids: oriented CSR (rows sorted), undirected CSR with
// the oriented edge id of every slot, supports and peel state per edge.
struct CoreView {
  uint32_t nc = 0, mc = 0, max_dout = 0;
  uint32_t *inv, *off_d, *adj_d, *src, *off_u, *adj_u, *eid_u, *state;
  int32_t* sup;
};

struct TrussOutcome {
  uint32_t kmax = 0;
  uint64_t count = 0;
  uint32_t stamp_floor = 0;  // survivors: state == kNone or state > floor
  uint32_t levels = 0;
  uint64_t rounds = 0;
};

class Solver {
 public:
  explicit Solver(int device);
  ~Solver();

  // Input already resident on the device: m canonical pairs (interleaved).
  void solve_pairs_device(const uint32_t* d_pairs, uint64_t m, uint32_t n, kt_result* res);
  // Directed line list (x into row y) resident on the device.
  void solve_lines_device(const uint32_t* d_xs, const uint32_t* d_ys, uint64_t L, uint32_t n,
                          kt_result* res);
  // Copies the survivors of the last solve (original ids, u < v) to host.
  uint64_t truss_edges(uint32_t* host_pairs, uint64_t cap);

  // Byte-model terms of the last pass-2 core.
  void last_stats(kt_stats* out);
  // Stage access for parity tests: device CSR of the last solve.
  void csr_to_host(uint64_t* off, uint32_t* adj);
  void set_lanes(int count_lanes, int peel_lanes) { lanes_count_ = count_lanes; lanes_peel_ = peel_lanes; }
  void set_force_edge_count(bool v) { force_edge_count_ = v; }

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

  // stage-level entry points (used by the reference-compatible API)
  void load_csr_host(uint32_t n, const uint64_t* off, const uint32_t* adj);
  uint64_t peel_vertices(uint32_t k);
  uint32_t max_core(uint64_t alive, std::vector<uint32_t>* ok_thresholds);
  void extract_core(CoreView* cv);
  void count_support(CoreView* cv, float* ms);
  TrussOutcome truss_search(CoreView* cv, int32_t thr0, uint64_t early_ans, uint32_t early_k);

  // state of the vertex peel (device)
  uint32_t* lvl() { return lvl_.as<uint32_t>(); }
  uint32_t* deg() { return deg_.as<uint32_t>(); }
  uint64_t* off() { return off_.as<uint64_t>(); }
  uint32_t* adj() { return adj_.as<uint32_t>(); }
  uint32_t n() const { return n_; }

 private:
  void build_from_pairs(const uint32_t* d_pairs, uint64_t m, uint32_t n);
  void build_from_lines(const uint32_t* d_xs, const uint32_t* d_ys, uint64_t L, uint32_t n);
  void solve_csr(kt_result* res);
  uint32_t read_u32(const uint32_t* dptr);
  void zero_u32(uint32_t* dptr);
  void exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint32_t n_plus_1);
  void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n_plus_1);
  uint32_t reduce_max_u32(const uint32_t* in, uint32_t n);

  int device_;
  hipStream_t stream_ = nullptr;
  Ctl* ctl_ = nullptr;
  uint32_t* pinned_ = nullptr;
  hipEvent_t ev_[8];
  int lanes_count_ = 0, lanes_peel_ = 0;
  bool force_edge_count_ = false;

  uint32_t n_ = 0;
  uint64_t m_ = 0;
  // full graph
  DevBuf off_, adj_, deg_, lvl_, q0_, q1_, deg_bak_, cursor_;
  // core
  DevBuf keys_, keys2_, to_, inv_, dout_, off_d_, adj_d_, src_, din_, din_off_, tkeys_, tvals_,
      tiota_, in_src_, off_u_, adj_u_, eid_u_, sup_, state_, eq0_, eq1_, out_pairs_;
  DevBuf temp_;
  CoreView last_core_{};
  uint32_t last_floor_ = 0;
  uint64_t last_truss_ = 0;
  bool have_truss_ = false;
};

}  // namespace kt
