// Device-resident k-max-truss solve: context, buffers, primitives and the
// (a) CSR build. The stage logic lives in kt_stages.hip.
#include "kt_solver.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "kt_kernels.h"

namespace kt {

namespace {
thread_local Arena* tl_arena = nullptr;
}  // namespace

ArenaScope::ArenaScope(Arena* a) : prev(tl_arena) { tl_arena = a; }
ArenaScope::~ArenaScope() { tl_arena = prev; }

void DevBuf::ensure(size_t b) {
  if (b <= bytes && p) return;
  if (Arena* a = tl_arena) {
    // carve: blocks are 256-byte aligned; the block on top of the arena
    // grows in place, any other is left behind (reset() reclaims it)
    const size_t want = (std::max<size_t>(b, 256) + 255) & ~size_t(255);
    uint8_t* top = a->base + a->used;
    if (p && carved && static_cast<uint8_t*>(p) + bytes == top && a->used - bytes + want <= a->size) {
      a->used += want - bytes;
      bytes = want;
      return;
    }
    if (a->used + want > a->size)
      fail("scratch exhausted: the solve needs more than " + std::to_string(a->size) +
           " bytes of device scratch (" + std::to_string(a->used) + " in use, " + std::to_string(want) +
           " more asked)");
    p = top;
    a->used += want;
    bytes = want;
    carved = true;
    return;
  }
  if (carved) fail("a buffer carved from scratch was grown outside its scratch scope");
  // queued work may still read the old block
  if (p) KT_HIP(hipDeviceSynchronize());
  release();
  size_t want = std::max<size_t>(b, 256);
  want = (want + 4095) & ~size_t(4095);
  KT_HIP(hipMalloc(&p, want));
  bytes = want;
}

void DevBuf::release() {
  if (p && !carved) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  carved = false;
}

namespace {
struct CastU64 {
  __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};
}  // namespace

Solver::Solver(int device) : device_(device) {
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  KT_HIP(hipMalloc(&ctl_, sizeof(Ctl)));
  KT_HIP(hipMemsetAsync(ctl_, 0, sizeof(Ctl), stream_));
  // 256 B of control-block and counter readbacks, then the degree histogram
  KT_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned_), 256 + sizeof(uint32_t) * kDegBins,
                       hipHostMallocDefault));
  pinned_hist_ = pinned_ + 64;
  init_common();
}

Solver::Solver(int device, Arena* arena) : device_(device), arena_(arena) {
  static_assert(sizeof(Ctl) <= 4096 && sizeof(Ctl) % 4 == 0, "control block exceeds its arena slot");
  if (!arena || arena->size < 8192 || (reinterpret_cast<uintptr_t>(arena->base) & 255))
    fail("arena mode: the arena must be 256-byte aligned and hold at least 8 KiB");
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  ctl_ = reinterpret_cast<Ctl*>(arena->base);  // the arena's first 4 KiB, kept across resets
  arena->floor = 4096;
  arena->reset();
  fill_u32(reinterpret_cast<uint32_t*>(ctl_), sizeof(Ctl) / 4, 0, stream_);
  pinned_ = host_rb_;
  init_common();
}

void Solver::forget_buffers() {
  each_buffer([&](const DevBuf& b) {
    if (b.p && !b.carved) fail("arena mode: a device buffer was allocated outside the arena");
    DevBuf& w = const_cast<DevBuf&>(b);
    w.p = nullptr;
    w.bytes = 0;
    w.carved = false;
  });
  have_truss_ = truss_ready_ = triples_ready_ = core_valid_ = false;
  last_core_ = CoreView{};
}

double Solver::trace_ms() {
  KT_HIP(hipStreamSynchronize(stream_));
  const auto now = std::chrono::steady_clock::now();
  const double ms = std::chrono::duration<double, std::milli>(now - trace_t0_).count();
  trace_t0_ = now;
  return ms;
}

void Solver::clear_control() { fill_u32(reinterpret_cast<uint32_t*>(ctl_), sizeof(Ctl) / 4, 0, stream_); }

void Solver::init_common() {
  static_assert(sizeof(host_rb_) >= 256, "readback buffer smaller than the pinned one");
  for (auto& e : ev_) KT_HIP(hipEventCreate(&e));
  for (auto& e : stage_ev_) KT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  KT_HIP(hipStreamSynchronize(stream_));
  // cost-model knobs of the frontier-versus-recount choice (tuning only)
  if (const char* v = std::getenv("KT_DENSE_TRI_COST")) dense_tri_cost_ = std::atof(v);
  if (const char* v = std::getenv("KT_DENSE_DROP_FIXED")) dense_drop_fixed_ = std::atof(v);
  if (const char* v = std::getenv("KT_SPARSE_DROP_RATIO")) sparse_drop_ratio_ = std::atof(v);
  if (const char* v = std::getenv("KT_TRACE")) trace_ = v[0] == '1';
  if (const char* v = std::getenv("KT_WALK_WINDOW")) walk_window_ = std::max<int64_t>(1, std::atoll(v));
  if (const char* v = std::getenv("KT_GALLOP_GROWTH")) gallop_growth_ = std::max<int64_t>(2, std::atoll(v));
  if (const char* v = std::getenv("KT_EIDM_BUDGET_MB")) eidm_budget_ = uint64_t(std::atoll(v)) << 20;
  if (const char* v = std::getenv("KT_WALK_LEVELS")) walk_levels_ = std::max(1, std::atoi(v));
  if (const char* v = std::getenv("KT_BUILD_TILES")) build_tiles_ = uint32_t(std::max(1, std::atoi(v)));
  if (const char* v = std::getenv("KT_VERTEX_PEEL"))
    vertex_mode_ = std::strcmp(v, "rounds") == 0 ? 1 : std::strcmp(v, "cascade") == 0 ? 2 : 0;
  default_drop_ratio_ = sparse_drop_ratio_;
}

template <typename F>
void Solver::each_buffer(F f) const {
  const DevBuf* all[] = {&off_, &adj_, &deg_, &lvl_, &q0_, &q1_, &deg_bak_, &cursor_, &list_, &new_id_,
                         &keys_, &keys2_, &vals_, &vals2_, &to_, &inv_, &inv_orig_, &dout_, &off_d_,
                         &adj_d_, &src_, &din_, &din_off_, &tkeys_, &tvals_, &tiota_, &in_src_, &off_u_,
                         &adj_u_, &eid_u_, &sup_, &state_, &eq0_, &eq1_, &out_pairs_, &truss_out_, &del_thr_, &orig0_,
                         &temp_, &hist_, &hoff_, &part_, &bits0_, &bits_, &fbits_, &pref0_, &eidm_, &heavy_,
                         &vflag_, &vmap_, &roff_, &dhist_, &vq_, &dec_out_, &bk_state_, &bk_sup_, &bk_bits_, &bk_del_, &text_,
                         &tile_cnt_, &tile_off_, &lx_, &ly_, &lkind_, &lx2_, &ly2_, &lkeep_, &lkeep_off_,
                         &lstat_, &probe_stats_};
  for (const DevBuf* b : all) f(*b);
  for (auto& w : work_)
    for (const DevBuf* b : {&w.off, &w.adj, &w.src, &w.orig, &w.sup}) f(*b);
  for (auto& l : levels_)
    for (const DevBuf* b : {&l->b_off, &l->b_adj, &l->b_deg, &l->b_lvl, &l->b_orig}) f(*b);
}

uint64_t Solver::device_bytes() const {
  uint64_t total = sizeof(Ctl);
  each_buffer([&](const DevBuf& b) { total += b.bytes; });
  return total;
}

void Solver::release_buffers() {
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipStreamSynchronize(stream_));
  each_buffer([](const DevBuf& b) { const_cast<DevBuf&>(b).release(); });
}

Solver::~Solver() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  each_buffer([](const DevBuf& b) { const_cast<DevBuf&>(b).release(); });
  for (auto& e : ev_) (void)hipEventDestroy(e);
  for (auto& e : stage_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : probe_ev_) (void)hipEventDestroy(e);
  for (auto* p : stage_)
    if (p) (void)hipHostFree(p);
  if (!arena_) {
    if (ctl_) (void)hipFree(ctl_);
    if (pinned_) (void)hipHostFree(pinned_);
  }
  if (stream_) (void)hipStreamDestroy(stream_);
}

uint32_t Solver::read_u32(const uint32_t* dptr) {
  KT_HIP(hipMemcpyAsync(pinned_, dptr, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  return pinned_[0];
}

uint64_t Solver::read_u64(const uint64_t* dptr) {
  KT_HIP(hipMemcpyAsync(pinned_, dptr, sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  uint64_t v;
  std::memcpy(&v, pinned_, sizeof(v));
  return v;
}

Ctl Solver::read_ctl() {
  static_assert(sizeof(Ctl) <= 256, "control block exceeds the pinned readback buffer");
  KT_HIP(hipMemcpyAsync(pinned_, ctl_, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  Ctl c;
  std::memcpy(&c, pinned_, sizeof(c));
  return c;
}

void Solver::zero_u32(uint32_t* dptr) { fill_u32(dptr, 1, 0, stream_); }

void Solver::exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint32_t count) {
  hipcub::TransformInputIterator<uint64_t, CastU64, const uint32_t*> it(in, CastU64());
  size_t tb = 0;
  KT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, it, out, count, stream_));
  temp_.ensure(tb);
  KT_HIP(hipcub::DeviceScan::ExclusiveSum(temp_.p, tb, it, out, count, stream_));
}

void Solver::exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t count) {
  size_t tb = 0;
  KT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, count, stream_));
  temp_.ensure(tb);
  KT_HIP(hipcub::DeviceScan::ExclusiveSum(temp_.p, tb, in, out, count, stream_));
}

uint32_t Solver::reduce_max_u32(const uint32_t* in, uint32_t count) {
  if (count == 0) return 0;
  uint32_t* out = &ctl_->pad[0];
  size_t tb = 0;
  KT_HIP(hipcub::DeviceReduce::Max(nullptr, tb, in, out, count, stream_));
  temp_.ensure(tb);
  KT_HIP(hipcub::DeviceReduce::Max(temp_.p, tb, in, out, count, stream_));
  return read_u32(out);
}

void Solver::sort_pairs_u64(const uint64_t* k_in, uint64_t* k_out, const uint32_t* v_in, uint32_t* v_out,
                            uint32_t count) {
  if (!count) return;
  size_t tb = 0;
  KT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k_in, k_out, v_in, v_out, count, 0, 64, stream_));
  temp_.ensure(tb);
  KT_HIP(hipcub::DeviceRadixSort::SortPairs(temp_.p, tb, k_in, k_out, v_in, v_out, count, 0, 64, stream_));
}

// ------------------------------------------------------------- (a) CSR ---

// Bucketed build (see k_part_hist / k_part_scatter / k_bucket_rows): random
// traffic is confined to one bucket's L2-resident slice of adj. Returns false
// when n is too large for the LDS row counters (then the atomic build runs).
bool Solver::build_partitioned(int mode, const uint32_t* a, const uint32_t* b, uint64_t items, uint32_t n) {
  uint32_t S = 0, B = 0;
  if (!csr_partition_plan(n, &S, &B)) return false;
  // pairs sorted by their first id: the first entries of a bucket's rows lie
  // in order in the input, so only the second entries are bucketed
  bool fwd_sorted = false;
  if (mode == 0 && sorted_build_ && items > 1) {
    zero_u32(&ctl_->pad[0]);
    csr_sorted_check(0, a, items, n, nullptr, nullptr, &ctl_->pad[0], &ctl_->err, stream_);
    fwd_sorted = read_u32(&ctl_->pad[0]) == 0;
  }
  const int pmode = fwd_sorted ? 2 : mode;
  const uint64_t entries = mode == 0 && !fwd_sorted ? 2 * items : items;  // bucketed entries
  // at most build_tiles_ tiles of at least 65536 items
  const uint64_t tiles = std::max<uint32_t>(1, build_tiles_);
  const uint64_t tile = std::max<uint64_t>(65536, (items + tiles - 1) / tiles);
  const uint32_t T = uint32_t(std::max<uint64_t>(1, (items + tile - 1) / tile));
  const uint64_t bt = uint64_t(B) * T;
  hist_.ensure(sizeof(uint32_t) * (2 * bt + 1));   // tile-major counts, then bucket-major
  hoff_.ensure(sizeof(uint64_t) * (2 * bt + 1));   // bucket-major offsets, then tile-major
  part_.ensure(sizeof(uint32_t) * 2 * std::max<uint64_t>(entries, 1));
  uint32_t* hist_t = hist_.as<uint32_t>();
  uint32_t* hist_b = hist_t + bt;
  uint64_t* hoff_b = hoff_.as<uint64_t>();
  uint64_t* hoff_t = hoff_b + bt + 1;
  uint64_t* f0 = fwd_sorted ? cursor_.as<uint64_t>() : nullptr;  // B + 1 <= n + 1 words
  fill_u32(hist_b + bt, 1, 0, stream_);
  zero_u32(&ctl_->err);
  csr_part_hist(pmode, a, b, items, tile, n, S, B, T, hist_t, &ctl_->err, stream_);
  if (read_u32(&ctl_->err)) fail("edge list holds a vertex id >= n");
  transpose_u32(hist_t, T, B, hist_b, stream_);
  exclusive_scan_u32_to_u64(hist_b, hoff_b, uint32_t(bt + 1));
  transpose_u64(hoff_b, B, T, hoff_t, stream_);
  if (fwd_sorted) csr_fwd_start(a, items, S, B, f0, stream_);
  csr_part_scatter(pmode, a, b, items, tile, n, S, B, T, hoff_t, part_.as<uint32_t>(), stream_);
  csr_bucket_rows(part_.as<uint32_t>(), hoff_b, fwd_sorted ? a : nullptr, f0, B, T, S, n, off(), deg(), adj(),
                  stream_);
  if (fwd_sorted)
    fill_u64(off() + n, 1, 2 * items, stream_);
  else
    KT_HIP(hipMemcpyAsync(off() + n, hoff_b + bt, sizeof(uint64_t), hipMemcpyDeviceToDevice, stream_));
  return true;
}

// Lines sorted by row (the reference's edge-list order): each row is a run
// of the lines, so the offsets come from the runs' ends and adj is xs as it
// is. Returns false (nothing built) when the rows are not sorted.
bool Solver::build_sorted_lines(const uint32_t* xs, const uint32_t* ys, uint64_t L, uint32_t n) {
  uint64_t* hpos = cursor_.as<uint64_t>();
  uint64_t* tpos = off();
  fill_u64(hpos, uint64_t(n) + 1, 0, stream_);
  fill_u64(tpos, uint64_t(n) + 1, 0, stream_);
  zero_u32(&ctl_->pad[0]);
  zero_u32(&ctl_->err);
  csr_sorted_check(1, ys, L, n, hpos, tpos, &ctl_->pad[0], &ctl_->err, stream_);
  if (read_u32(&ctl_->pad[0]) != 0) return false;
  if (read_u32(&ctl_->err)) fail("edge list holds a vertex id >= n");
  csr_deg_from_runs(hpos, tpos, n, deg(), stream_);
  exclusive_scan_u32_to_u64(deg(), off(), n + 1);
  csr_copy_cols(xs, L, n, adj(), &ctl_->err, stream_);
  if (read_u32(&ctl_->err)) fail("edge list holds a vertex id >= n");
  return true;
}

void Solver::new_input() {
  deg_at_least_.clear();
  have_truss_ = false;
  truss_ready_ = false;
  triples_ready_ = false;
}

void Solver::build_from_pairs(const uint32_t* d_pairs, uint64_t m, uint32_t n) {
  new_input();
  if (reinterpret_cast<uintptr_t>(d_pairs) % 8 != 0) fail("pairs must be 8-byte aligned");
  if (2 * m >= (uint64_t(1) << 40)) fail("edge list too large");
  n_ = n;
  m_ = m;
  deg_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  off_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  cursor_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  adj_.ensure(sizeof(uint32_t) * std::max<uint64_t>(2 * m, 1));
  if (!force_atomic_build_ && build_partitioned(0, d_pairs, nullptr, m, n)) return;
  fill_u32(deg(), uint64_t(n) + 1, 0, stream_);
  zero_u32(&ctl_->err);
  csr_degree_pairs(d_pairs, m, n, deg(), ctl_, stream_);
  exclusive_scan_u32_to_u64(deg(), off(), n + 1);
  KT_HIP(hipMemcpyAsync(cursor_.p, off_.p, sizeof(uint64_t) * (size_t(n) + 1),
                        hipMemcpyDeviceToDevice, stream_));
  csr_scatter_pairs(d_pairs, m, n, cursor_.as<uint64_t>(), adj(), stream_);
  if (read_u32(&ctl_->err)) fail("edge list holds a vertex id >= n");
}

void Solver::build_from_lines(const uint32_t* d_xs, const uint32_t* d_ys, uint64_t L, uint32_t n) {
  new_input();
  n_ = n;
  m_ = L / 2;
  deg_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  off_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  cursor_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  adj_.ensure(sizeof(uint32_t) * std::max<uint64_t>(L, 1));
  if (!force_atomic_build_ && sorted_build_ && build_sorted_lines(d_xs, d_ys, L, n)) return;
  if (!force_atomic_build_ && build_partitioned(1, d_xs, d_ys, L, n)) return;
  fill_u32(deg(), uint64_t(n) + 1, 0, stream_);
  zero_u32(&ctl_->err);
  csr_degree_lines(d_xs, d_ys, L, n, deg(), ctl_, stream_);
  exclusive_scan_u32_to_u64(deg(), off(), n + 1);
  KT_HIP(hipMemcpyAsync(cursor_.p, off_.p, sizeof(uint64_t) * (size_t(n) + 1),
                        hipMemcpyDeviceToDevice, stream_));
  csr_scatter_lines(d_xs, d_ys, L, n, cursor_.as<uint64_t>(), adj(), stream_);
  if (read_u32(&ctl_->err)) fail("edge list holds a vertex id >= n");
}

void Solver::load_csr_host(uint32_t n, const uint64_t* off, const uint32_t* adj) {
  new_input();
  n_ = n;
  m_ = off[n] / 2;
  deg_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  off_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  adj_.ensure(sizeof(uint32_t) * std::max<uint64_t>(off[n], 1));
  KT_HIP(hipMemcpyAsync(off_.p, off, sizeof(uint64_t) * (size_t(n) + 1), hipMemcpyHostToDevice, stream_));
  if (off[n])
    KT_HIP(hipMemcpyAsync(adj_.p, adj, sizeof(uint32_t) * off[n], hipMemcpyHostToDevice, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
}

void Solver::csr_to_host(uint64_t* off, uint32_t* adj) {
  KT_HIP(hipMemcpyAsync(off, off_.p, sizeof(uint64_t) * (size_t(n_) + 1), hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  if (off[n_])
    KT_HIP(hipMemcpyAsync(adj, adj_.p, sizeof(uint32_t) * off[n_], hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
}

LevelGraph& Solver::level0() {
  if (levels_.empty()) levels_.emplace_back(new LevelGraph);
  LevelGraph& g = *levels_[0];
  lvl_.ensure(sizeof(uint32_t) * (size_t(n_) + 1));
  g.n = n_;
  g.slots = read_u64(off() + n_);
  g.off = off();
  g.adj = adj();
  g.deg = deg();
  g.lvl = lvl_.as<uint32_t>();
  g.to_orig = nullptr;
  g.k_enter = 0;
  g.ok.clear();
  g.exact_from = g.exact_to = 0;
  return g;
}

hipEvent_t* Solver::probe_events(int launches) {
  while (2 * (probe_used_ + size_t(launches)) > probe_ev_.size())
    for (int e = 0; e < 64; ++e) {
      probe_ev_.push_back(nullptr);
      KT_HIP(hipEventCreate(&probe_ev_.back()));
    }
  hipEvent_t* ev = probe_ev_.data() + 2 * probe_used_;
  probe_used_ += size_t(launches);
  return ev;
}

void Solver::set_probe(int kernel, int mode) {
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipStreamSynchronize(stream_));
  if (kernel != kProbeVertexPeel && kernel != kProbeDensePeel) fail("probe: unknown kernel");
  probe_kernel_ = kernel;
  probe_push_ = mode & 3;
  probe_used_ = 0;
  if (probe_push_) {
    probe_stats_.ensure(2 * sizeof(unsigned long long));
    fill_u64(probe_stats_.as<uint64_t>(), 2, 0, stream_);
  }
}

void Solver::push_stats(uint64_t* launches, double* ms, uint64_t* slots, uint64_t* vertices) {
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipStreamSynchronize(stream_));
  double total = 0;
  for (size_t i = 0; i < probe_used_; ++i) {
    float t = 0;
    KT_HIP(hipEventElapsedTime(&t, probe_ev_[2 * i], probe_ev_[2 * i + 1]));
    total += t;
  }
  uint64_t h[2] = {0, 0};
  if (probe_stats_.p) KT_HIP(hipMemcpy(h, probe_stats_.p, sizeof(h), hipMemcpyDeviceToHost));
  *launches = probe_used_;
  *ms = total;
  *slots = h[0];
  *vertices = h[1];
}

}  // namespace kt
