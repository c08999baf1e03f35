// Host orchestration of the device-resident k-max-truss solve.
#pragma once
#include <chrono>
#include <memory>
#include <vector>

#include "kt_common.h"
#include "../../include/ktruss.h"

namespace kt {

// A fixed block of device memory that buffers are carved from, in bump
// order, instead of being allocated. reset() frees everything above `floor`
// at once; nothing is ever returned to the HIP allocator.
struct Arena {
  uint8_t* base = nullptr;
  size_t size = 0;
  size_t floor = 0;  // bytes that survive reset()
  size_t used = 0;
  void reset() { used = floor; }
};

// While an ArenaScope is alive, DevBuf::ensure on the same host thread
// carves from its arena and never calls hipMalloc (and DevBuf::release of a
// carved buffer frees nothing).
struct ArenaScope {
  explicit ArenaScope(Arena* a);
  ~ArenaScope();
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;
  Arena* prev;
};

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool carved = false;  // p lies in an arena
  void ensure(size_t b);
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// One graph of the vertex-peel nest. Level 0 is the input CSR; each deeper
// level is the induced subgraph of the vertices still alive when it was cut,
// with its own dense ids (to_orig maps them back to input ids).
struct LevelGraph {
  uint32_t n = 0;
  uint64_t slots = 0;
  uint64_t* off = nullptr;
  uint32_t* adj = nullptr;
  uint32_t* deg = nullptr;       // alive-neighbour count of alive vertices
  uint32_t* lvl = nullptr;       // kNone: alive; else threshold that removed it
  uint32_t* to_orig = nullptr;   // null on level 0
  uint32_t k_enter = 0;          // every vertex here has core number >= k_enter
  std::vector<uint32_t> ok;      // thresholds whose peel left vertices alive
  // after a core decomposition: lvl is exact for thresholds in
  // [exact_from, exact_to] (the t-core = alive or lvl > t); 0, 0 otherwise
  uint32_t exact_from = 0, exact_to = 0;
  DevBuf b_off, b_adj, b_deg, b_lvl, b_orig;
};

// Core subgraph in rank ids: oriented CSR (rows sorted), undirected CSR with
// the oriented edge id of every slot, supports and peel state per edge.
// The base edge set (the extraction's output) is never modified; compaction
// writes the surviving edges into one of two work sets and `orig` maps each
// current edge back to its base id, where its deletion threshold is recorded.
struct CoreView {
  uint32_t nc = 0, mc = 0, max_dout = 0;
  uint32_t nc0 = 0;  // vertices of the base set (rank ids index inv)
  uint32_t *inv = nullptr, *off_d = nullptr, *adj_d = nullptr, *src = nullptr, *orig = nullptr;
  uint32_t *off_u = nullptr, *adj_u = nullptr, *eid_u = nullptr, *state = nullptr;
  int32_t* sup = nullptr;
  // base set
  uint32_t mc0 = 0, max_dout0 = 0;
  uint32_t *off_d0 = nullptr, *adj_d0 = nullptr, *src0 = nullptr, *del_thr = nullptr;
  int work = -1;  // which work set holds the current edges (-1: base)
  // dense cores: bit-matrix rows (ws words each) of all / live / frontier
  // edges, per-word prefix popcounts of the static rows, and (when nc^2
  // entries fit the budget) the edge id of every (u, w) pair
  bool dense = false;
  uint32_t ws = 0;
  uint64_t *bits0 = nullptr, *bits = nullptr, *fbits = nullptr;
  uint32_t* pref0 = nullptr;
  uint32_t* eidm = nullptr;
};

struct TrussOutcome {
  uint32_t kmax = 0;
  uint64_t count = 0;
  uint64_t alive = 0;      // live edges when the search stopped
  bool cut_short = false;  // stopped by the level budget with live edges left
  int32_t floor_thr = -1;  // survivors: del_thr == kNone or del_thr > floor_thr
  uint32_t levels = 0;
  uint64_t rounds = 0;
  uint32_t compactions = 0;  // edge sets rewritten (compaction, re-densify)
  uint32_t recounts = 0;     // dropped frontiers followed by an exact recount
};

struct EdgeSet {
  DevBuf off, adj, src, orig, sup;
};

class Solver {
 public:
  explicit Solver(int device);
  // Arena mode: the control block and, inside an ArenaScope on `arena`,
  // every device buffer come from the arena; readbacks go through host
  // memory of the object itself. Nothing is allocated on the device.
  Solver(int device, Arena* arena);
  ~Solver();
  // Arena mode: drops every buffer (all carved) before the arena is reset.
  void forget_buffers();
  // Zeroes the control block (a fill kernel on the solver's stream).
  void clear_control();
  // Device bytes this context holds now (its buffers keep their size
  // between solves, so after a solve this is at least that solve's peak).
  uint64_t device_bytes() const;
  // Frees every buffer (they grow again on the next solve); the last solve's
  // device-side results are gone afterwards.
  void release_buffers();
  // Largest edge-id matrix a dense core may allocate (bytes).
  void set_eidm_budget(uint64_t bytes) { eidm_budget_ = bytes; }

  // Input already resident on the device: m canonical pairs (interleaved).
  void solve_pairs_device(const uint32_t* d_pairs, uint64_t m, uint32_t n, kt_result* res);
  // Truss decomposition of the 2-core (a full level walk from threshold 1:
  // no max-core pass, no early exit). On return *cv holds the base core
  // edges (src0 / adj_d0 in rank ids, inv: rank -> input vertex id) and
  // del_thr: an edge deleted at threshold t has truss number t + 1, an edge
  // never deleted has the returned kmax (0 when the 2-core has no edges).
  uint32_t decompose_pairs_device(const uint32_t* d_pairs, uint64_t m, uint32_t n, CoreView* cv);
  // Text edge list (the reference's contract) read into HBM and parsed
  // there, then solved; false (nothing solved) when a line holds a single
  // number, which the host parser handles. parse_ms: file to line arrays.
  bool solve_tsv_device(const char* path, kt_result* res, double* parse_ms);
  // The parse alone; the lines stay on the device (copy_ingested reads them).
  bool ingest_tsv(const char* path, uint64_t* lines, uint32_t* n);
  void copy_ingested(uint32_t* xs, uint32_t* ys, uint64_t count);
  // Directed line list (x into row y) resident on the device.
  void solve_lines_device(const uint32_t* d_xs, const uint32_t* d_ys, uint64_t L, uint32_t n,
                          kt_result* res);
  // Copies the survivors of the last solve (original ids, u < v) to host.
  uint64_t truss_edges(uint32_t* host_pairs, uint64_t cap);
  // The same into device memory (queued on the solver's stream, not waited for).
  uint64_t truss_edges_to_device(uint32_t* dev_pairs, uint64_t cap);
  // Truss decomposition of the last input graph: (u, v, t) with u < v for
  // every edge of trussness t >= 3 (the others have 2), unordered; returns
  // the count, copies at most cap triples. Replaces the last solve's
  // survivors (truss_edges fails afterwards).
  uint64_t trussness(uint32_t* host_triples, uint64_t cap);
  // Byte-model terms of the last pass-2 core.
  void last_stats(kt_stats* out);
  // Stage access for parity tests: device CSR of the last solve.
  void csr_to_host(uint64_t* off, uint32_t* adj);
  void set_lanes(int count_lanes, int peel_lanes) { lanes_count_ = count_lanes; lanes_peel_ = peel_lanes; }
  // Measurement of the vertex-peel push kernel: mode bit 1 stamps events at
  // each launch's start and end, bit 2 counts the light row slots it walks
  // and the vertices it takes (separate solves keep the counters' atomics
  // out of the timing); push_stats sums them since set_push_probe.
  void set_push_probe(int mode) { set_probe(kProbeVertexPeel, mode); }
  // kernel: kProbeVertexPeel (the vertex-peel cascade, or the push of its
  // rounds) or kProbeDensePeel (the bit-matrix edge-peel round); for the
  // latter the counters are the frontier edges' d+(u) + d+(v) and their
  // number.
  static constexpr int kProbeVertexPeel = 0;
  static constexpr int kProbeDensePeel = 1;
  void set_probe(int kernel, int mode);
  void push_stats(uint64_t* launches, double* ms, uint64_t* slots, uint64_t* vertices);
  // Settings the batch autotuner varies (every value computes the same result).
  struct Tuning {
    int count_lanes = 0, peel_lanes = 0;  // 0: by average degree
    bool edge_count = false;              // per-edge support kernel
    double drop_ratio = 3.0;              // sparse peel: recount when frontier * ratio >= live edges
  };
  Tuning tuning() const { return Tuning{lanes_count_, lanes_peel_, force_edge_count_, sparse_drop_ratio_}; }
  void set_tuning(const Tuning& t) {
    lanes_count_ = t.count_lanes;
    lanes_peel_ = t.peel_lanes;
    force_edge_count_ = t.edge_count;
    sparse_drop_ratio_ = t.drop_ratio;
  }
  void reset_drop_ratio() { sparse_drop_ratio_ = default_drop_ratio_; }
  void set_force_edge_count(bool v) { force_edge_count_ = v; }
  void set_force_atomic_build(bool v) { force_atomic_build_ = v; }
  void set_sorted_build(bool v) { sorted_build_ = v; }
  void set_core_mode(int m) { core_mode_ = m; }  // 0 auto, 1 sparse, 2 dense
  // pass 1 on dense views: walk `walk_levels` levels, then probe thresholds
  void set_gallop(bool v, uint32_t walk_levels) {
    gallop_ = v;
    walk_levels_ = walk_levels ? walk_levels : 1;
  }

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

  // stages (kt_stages.hip)
  void load_csr_host(uint32_t n, const uint64_t* off, const uint32_t* adj);
  uint64_t peel_vertices(LevelGraph& g, uint32_t k, uint64_t alive = 0);
  uint64_t peel_vertices_rounds(LevelGraph& g, uint32_t k, uint64_t stop_at, uint64_t alive);
  // vertex peel: 0 by level size (one cascade launch per threshold up to
  // kCascadeMaxSlots row slots, synchronous rounds above), 1 always rounds,
  // 2 always the cascade
  void set_vertex_mode(int m) { vertex_mode_ = m; }
  static constexpr uint64_t kCascadeMaxSlots = uint64_t(1) << 24;
  // a level of at most core_decompose_max_n() vertices and this many row
  // slots finishes the max-core search in one workgroup (decompose_level)
  // when its rows average at most kDecomposeMaxRow slots or at most
  // kDecomposeMaxRange thresholds are left
  static constexpr uint64_t kDecomposeMaxSlots = uint64_t(1) << 25;
  static constexpr uint32_t kDecomposeMaxRow = 2048;
  static constexpr uint32_t kDecomposeMaxRange = 64;
  // fresh: start from full degrees with every vertex alive; otherwise from
  // the state already on level 0. allow_cut: continue on induced subgraphs.
  uint32_t max_core(bool allow_cut = true, bool fresh = true);
  uint32_t decompose_level(LevelGraph& g, uint32_t k);
  bool filtered_start(LevelGraph** gp, uint32_t* k_out, uint64_t* alive_out);
  LevelGraph& level0();
  // recount: the alive vertices' degrees in `parent` are stale (a degree
  // filter stamped vertices without decrementing their neighbours), so the
  // cut counts them from the rows.
  LevelGraph& cut_level(LevelGraph& parent, uint32_t k_enter, uint32_t alive, bool recount = false);
  // The k-core of the input by a degree filter: vertices of degree < k are
  // stamped removed at k on level 0, the rest is cut (degrees recounted) and
  // peeled at k. Returns the k-core's vertex count; depth_ = 2 when it is
  // non-empty.
  uint64_t filtered_core(uint32_t k);
  // Vertices of input degree >= k (exact below kDegExact, else a lower
  // bound), from the degree histogram of the current input; ~0 when none.
  uint64_t degree_at_least(uint32_t k) const;
  void core_at(uint32_t k);  // leaves the k-core alive on the deepest usable level
  void extract_core(LevelGraph& g, CoreView* cv);
  void sort_oriented_rows(uint32_t nc, uint32_t mc, uint32_t max_row);
  void build_undirected_view(CoreView* cv);
  void count_support(CoreView* cv, float* ms);
  // Keep only the live edges (state == kNone) of the current view; with
  // carry_sup their supports are kept, otherwise recounted exactly. With a
  // vertex map (exclusive scan of live-vertex flags, nc2 kept vertices) the
  // vertices are renumbered too and a dense view is rebuilt at its new size.
  void compact_core(CoreView* cv, bool carry_sup, const uint32_t* vmap = nullptr, uint32_t nc2 = 0);
  // Dense views: after a dropped frontier, rebuild on the vertices that still
  // have live edges when that shrinks the rows enough; recounts supports.
  bool try_redensify(CoreView* cv);
  // exact_early: stop at a first level of early_ans edges and report early_k
  // whatever the minimum support (the reference's rule, src/kc.cu:1240).
  TrussOutcome truss_search(CoreView* cv, int32_t thr0, uint64_t early_ans, uint32_t early_k,
                            bool exact_early = false);
  // Dense views only: the same answer as truss_search(cv, 1, 0, 0) found by
  // probing thresholds with backups instead of walking every level.
  // empty_from: a threshold known to leave the view empty (-1: none)
  TrussOutcome truss_gallop(CoreView* cv, int64_t empty_from);

  // reference-compatible stage entry points (kt_compat.hip)
  void vertex_state_from_host(const uint32_t* deg_origin, const uint32_t* mask);
  void level0_mask_to_host(uint32_t* mask);
  void level0_degrees_to_host(uint32_t* deg);
  void orient_with_ranks(const uint32_t* h_to, uint32_t new_n, const uint32_t* h_off_d,
                         uint32_t* h_adj_d);
  void view_from_oriented_host(uint32_t nc, const uint32_t* h_off_d, const uint32_t* h_adj_d,
                               CoreView* cv);
  void undirected_to_host(const CoreView& cv, uint32_t* adj_u, uint32_t* eid_u, uint32_t* lson,
                          uint32_t* rson);
  void core_ranks_to_host(const CoreView& cv, uint32_t* inv_orig, uint32_t* dout);
  void support_to_host(const CoreView& cv, int32_t* sup);

  uint32_t* deg() { return deg_.as<uint32_t>(); }
  uint64_t* off() { return off_.as<uint64_t>(); }
  uint32_t* adj() { return adj_.as<uint32_t>(); }
  uint32_t n() const { return n_; }

 private:
  template <typename F>
  void each_buffer(F f) const;  // every device buffer the context owns
  void build_from_pairs(const uint32_t* d_pairs, uint64_t m, uint32_t n);
  void build_from_lines(const uint32_t* d_xs, const uint32_t* d_ys, uint64_t L, uint32_t n);
  bool build_partitioned(int mode, const uint32_t* a, const uint32_t* b, uint64_t items, uint32_t n);
  bool build_sorted_lines(const uint32_t* xs, const uint32_t* ys, uint64_t L, uint32_t n);
  void solve_csr(kt_result* res);
  uint32_t read_u32(const uint32_t* dptr);
  uint64_t read_u64(const uint64_t* dptr);
  Ctl read_ctl();
  RoundRule round_rule(const CoreView& cv) const;
  PeelArgs peel_args(const CoreView& cv);
  // frontier of the live edges below thr, stamped; the rounds' starting state
  // (reset: a new search with its early-exit rule)
  void start_rounds(CoreView* cv, int32_t thr, uint32_t r, uint64_t alive, bool reset, const SearchRule& search);
  TrussOutcome run_levels(CoreView* cv, int32_t thr, uint64_t alive, uint32_t* r_io, const SearchRule& search,
                          bool allow_redensify);
  void zero_u32(uint32_t* dptr);
  void exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint32_t n_plus_1);
  void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n_plus_1);
  uint32_t reduce_max_u32(const uint32_t* in, uint32_t n);
  void sort_pairs_u64(const uint64_t* k_in, uint64_t* k_out, const uint32_t* v_in, uint32_t* v_out,
                      uint32_t count);

  void init_common();
  int device_;
  Arena* arena_ = nullptr;  // arena mode
  hipStream_t stream_ = nullptr;
  Ctl* ctl_ = nullptr;
  uint32_t* pinned_ = nullptr;
  uint32_t* pinned_hist_ = nullptr;  // kDegBins words of pinned memory (none in arena mode)
  uint32_t host_rb_[64] = {};  // readback buffer in arena mode
  hipEvent_t ev_[10];
  int lanes_count_ = 0, lanes_peel_ = 0;
  bool force_edge_count_ = false;
  bool force_atomic_build_ = false;
  bool sorted_build_ = true;  // row-sorted inputs skip (part of) the scatter
  int64_t walk_window_ = 32;   // KT_WALK_WINDOW: pass-1 thresholds left when probes give way to walking
  int64_t gallop_growth_ = 2;  // KT_GALLOP_GROWTH: factor of the pass-1 probe steps

  int core_mode_ = 0;
  DevBuf bits0_, bits_, fbits_, pref0_, eidm_, heavy_, vflag_, vmap_, roff_, dhist_, vq_, dec_out_;
  int vertex_mode_ = 0;        // see set_vertex_mode
  bool vertex_rounds_ = false;  // the last vertex peel ran synchronous rounds (reported by last_stats)
  std::vector<uint64_t> deg_at_least_;  // input degree histogram, suffix sums per bin (empty: none)
  DevBuf bk_state_, bk_sup_, bk_bits_, bk_del_;  // threshold-probe backups
  uint64_t eidm_budget_ = uint64_t(4) << 30;  // bytes for the edge-id matrix
  double dense_tri_cost_ = 10.0, sparse_drop_ratio_ = 3.0, default_drop_ratio_ = 3.0;
  double dense_drop_fixed_ = 32e6;  // about what a drop's host round trip and restart cost
  bool trace_ = false;  // KT_TRACE=1: per-round frontier sizes on stderr
  // trace only: milliseconds since the previous call, after the stream drains
  double trace_ms();
  std::chrono::steady_clock::time_point trace_t0_{};
  bool gallop_ = true;
  uint32_t count_kind_ = 0, count_words_ = 0;  // the last timed support count
  uint32_t walk_levels_ = 16;
  uint32_t build_tiles_ = 4096;  // most tiles of the partitioned CSR build
  int probe_push_ = 0;               // measurement of a kernel (mode bits)
  int probe_kernel_ = kProbeVertexPeel;
  hipEvent_t* probe_events(int launches);  // 2 * launches events from the pool
  std::vector<hipEvent_t> probe_ev_;
  size_t probe_used_ = 0;
  DevBuf probe_stats_;
  void make_dense(CoreView* cv);
  HeavyList heavy() {
    heavy_.ensure(sizeof(uint32_t) * 2 * (size_t(n_) + 2));
    return HeavyList{heavy_.as<uint32_t>(), &ctl_->pad[1], heavy_.as<uint32_t>() + size_t(n_) + 2};
  }

  uint32_t n_ = 0;
  uint64_t m_ = 0;
  // full graph
  DevBuf off_, adj_, deg_, lvl_, q0_, q1_, deg_bak_, cursor_, hist_, hoff_, part_;
  std::vector<std::unique_ptr<LevelGraph>> levels_;
  size_t depth_ = 0;  // levels in use
  // scratch for level cuts and extraction
  DevBuf list_, new_id_, keys_, keys2_, vals_, vals2_, to_;
  // core
  DevBuf inv_, inv_orig_, dout_, off_d_, adj_d_, src_, din_, din_off_, tkeys_, tvals_, tiota_,
      in_src_, off_u_, adj_u_, eid_u_, sup_, state_, eq0_, eq1_, out_pairs_, del_thr_, orig0_;
  EdgeSet work_[2];
  DevBuf temp_;
  // device-side text ingest: bytes, per-tile line counts and offsets, lines
  DevBuf text_, tile_cnt_, tile_off_, lx_, ly_, lkind_, lx2_, ly2_, lkeep_, lkeep_off_, lstat_;
  uint8_t* stage_[2] = {nullptr, nullptr};  // pinned staging of the file
  hipEvent_t stage_ev_[2] = {nullptr, nullptr};
  CoreView last_core_{};
  int32_t last_floor_ = -1;
  uint64_t last_truss_ = 0;
  bool have_truss_ = false;    // a solve completed on the current input
  bool core_valid_ = false;    // last_core_ still describes that solve's core buffers
  bool truss_ready_ = false;   // its kmax-truss pairs are in truss_out_
  bool triples_ready_ = false; // the current input's trussness triples are in out_pairs_
  uint64_t truss_count_ = 0, triples_count_ = 0;
  DevBuf truss_out_;
  void new_input();            // a new graph: results of the previous one are gone
  void materialize_truss();
};

}  // namespace kt
