// Device-resident two-pass k-max-truss solve (host side).
//
// Flow (reference: src/kc_gpu.cpp:129-208, 336-341; src/kc.cu:770-884, 1140-1309):
//   (a) undirected CSR from the edge list, in HBM
//   (b) vertex peel with a galloping threshold -> max-core C0 (every removal
//       records its threshold in lvl[], so any smaller core is recovered later
//       without re-peeling the whole graph)
//   pass 1: rank + orient C0, exact supports, truss levels -> k̂ <= kmax
//   pass 2: the (L-1)-core C, L = max(k̂, 3) (it holds the kmax-truss),
//       rank + orient, exact supports, truss levels upward from L with the
//       reference's early exit (first level equal to pass 1's truss).
// Nothing returns to the host but scalar counters.
#include "kt_solver.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "kt_kernels.h"

namespace kt {

void DevBuf::ensure(size_t b) {
  if (b <= bytes && p) return;
  // queued work may still read the old block
  if (p) KT_HIP(hipDeviceSynchronize());
  release();
  size_t want = std::max<size_t>(b, 256);
  want = (want + 4095) & ~size_t(4095);
  KT_HIP(hipMalloc(&p, want));
  bytes = want;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

namespace {
struct CastU64 {
  __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

inline int auto_lanes(double avg) {
  if (avg >= 48) return 64;
  if (avg >= 20) return 32;
  if (avg >= 10) return 16;
  return 8;
}
}  // namespace

Solver::Solver(int device) : device_(device) {
  KT_HIP(hipSetDevice(device_));
  KT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  KT_HIP(hipMalloc(&ctl_, sizeof(Ctl)));
  KT_HIP(hipMemsetAsync(ctl_, 0, sizeof(Ctl), stream_));
  KT_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned_), 64, hipHostMallocDefault));
  for (auto& e : ev_) KT_HIP(hipEventCreate(&e));
  KT_HIP(hipStreamSynchronize(stream_));
}

Solver::~Solver() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  DevBuf* all[] = {&off_, &adj_, &deg_, &lvl_, &q0_, &q1_, &deg_bak_, &cursor_, &keys_, &keys2_,
                   &to_, &inv_, &dout_, &off_d_, &adj_d_, &src_, &din_, &din_off_, &tkeys_,
                   &tvals_, &tiota_, &in_src_, &off_u_, &adj_u_, &eid_u_, &sup_, &state_, &eq0_,
                   &eq1_, &out_pairs_, &temp_};
  for (auto* b : all) b->release();
  for (auto& e : ev_) (void)hipEventDestroy(e);
  if (ctl_) (void)hipFree(ctl_);
  if (pinned_) (void)hipHostFree(pinned_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

uint32_t Solver::read_u32(const uint32_t* dptr) {
  KT_HIP(hipMemcpyAsync(pinned_, dptr, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  return pinned_[0];
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

// ------------------------------------------------------------- (a) CSR ---

void Solver::build_from_pairs(const uint32_t* d_pairs, uint64_t m, uint32_t n) {
  if (reinterpret_cast<uintptr_t>(d_pairs) % 8 != 0) fail("pairs must be 8-byte aligned");
  n_ = n;
  m_ = m;
  deg_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  off_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  cursor_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  adj_.ensure(sizeof(uint32_t) * std::max<uint64_t>(2 * m, 1));
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
  n_ = n;
  m_ = L / 2;
  deg_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  off_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  cursor_.ensure(sizeof(uint64_t) * (size_t(n) + 1));
  adj_.ensure(sizeof(uint32_t) * std::max<uint64_t>(L, 1));
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

// ------------------------------------------------------ (b) vertex peel ---

// Removes every alive vertex of degree < k, round by round; returns the
// number removed. lvl[] of each removed vertex becomes k.
uint64_t Solver::peel_vertices(uint32_t k) {
  uint32_t* qc = ctl_->qcount;
  uint32_t* q[2] = {q0_.as<uint32_t>(), q1_.as<uint32_t>()};
  zero_u32(&qc[0]);
  core_frontier_init(lvl(), deg(), n_, k, q[0], &qc[0], stream_);
  uint32_t c = read_u32(&qc[0]);
  uint64_t removed = 0;
  int cur = 0;
  while (c) {
    core_stamp(q[cur], &qc[cur], c, k, lvl(), stream_);
    zero_u32(&qc[cur ^ 1]);
    core_process(off(), adj(), q[cur], &qc[cur], c, k, lvl(), deg(), q[cur ^ 1], &qc[cur ^ 1], stream_);
    removed += c;
    cur ^= 1;
    c = read_u32(&qc[cur]);
  }
  return removed;
}

// Galloping search for the largest k with a non-empty k-core, starting from
// the current state (all vertices of degree >= 1 alive). Leaves the max-core
// alive and returns its k. Successful thresholds are appended to
// ok_thresholds (removal thresholds recorded in lvl[] are among them).
uint32_t Solver::max_core(uint64_t alive, std::vector<uint32_t>* ok) {
  uint32_t k = 1, step = 16;
  bool grow = true;
  deg_bak_.ensure(sizeof(uint32_t) * (size_t(n_) + 1));
  while (step > 0) {
    const uint32_t kk = k + step;
    KT_HIP(hipMemcpyAsync(deg_bak_.p, deg_.p, sizeof(uint32_t) * n_, hipMemcpyDeviceToDevice, stream_));
    const uint64_t removed = peel_vertices(kk);
    if (removed < alive) {
      alive -= removed;
      k = kk;
      ok->push_back(k);
      step = grow ? step * 2 : step / 2;
    } else {
      // overshoot: re-open what this step removed and restore degrees
      core_reopen_above(lvl(), n_, k, stream_);
      KT_HIP(hipMemcpyAsync(deg_.p, deg_bak_.p, sizeof(uint32_t) * n_, hipMemcpyDeviceToDevice, stream_));
      if (step == 1) break;
      step /= 2;
      grow = false;
    }
  }
  return k;
}

// --------------------------------------------------- (a1-a3) extraction ---

void Solver::extract_core(CoreView* cv) {
  const uint32_t n = n_;
  keys_.ensure(sizeof(uint64_t) * std::max<uint32_t>(n, 1));
  keys2_.ensure(sizeof(uint64_t) * std::max<uint32_t>(n, 1));
  zero_u32(&ctl_->nalive);
  core_collect(lvl(), deg(), n, keys_.as<uint64_t>(), &ctl_->nalive, stream_);
  const uint32_t nc = read_u32(&ctl_->nalive);
  *cv = CoreView{};
  cv->nc = nc;
  to_.ensure(sizeof(uint32_t) * std::max<uint32_t>(n, 1));
  inv_.ensure(sizeof(uint32_t) * std::max<uint32_t>(nc, 1));
  dout_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  off_d_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  din_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  din_off_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  off_u_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  if (nc) {
    size_t tb = 0;
    KT_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys_.as<uint64_t>(), keys2_.as<uint64_t>(), nc,
                                             0, 64, stream_));
    temp_.ensure(tb);
    KT_HIP(hipcub::DeviceRadixSort::SortKeys(temp_.p, tb, keys_.as<uint64_t>(), keys2_.as<uint64_t>(), nc,
                                             0, 64, stream_));
  }
  fill_u32(to_.as<uint32_t>(), n, kNone, stream_);
  core_assign_rank(keys2_.as<uint64_t>(), nc, to_.as<uint32_t>(), inv_.as<uint32_t>(), stream_);
  fill_u32(dout_.as<uint32_t>(), uint64_t(nc) + 1, 0, stream_);
  orient_count(off(), adj(), to_.as<uint32_t>(), inv_.as<uint32_t>(), nc, dout_.as<uint32_t>(), stream_);
  // off_d sizes: the core of a graph with <= 2^32-1 directed slots has
  // m_C < 2^31; checked below before any use.
  exclusive_scan_u32(dout_.as<uint32_t>(), off_d_.as<uint32_t>(), nc + 1);
  const uint32_t mc = read_u32(off_d_.as<uint32_t>() + nc);
  if (mc >= 0x7FFFFFFFu) fail("core subgraph exceeds 2^31 edges");
  cv->mc = mc;
  cv->max_dout = reduce_max_u32(dout_.as<uint32_t>(), nc);
  adj_d_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  src_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  fill_u32(din_.as<uint32_t>(), uint64_t(nc) + 1, 0, stream_);
  orient_fill(off(), adj(), to_.as<uint32_t>(), inv_.as<uint32_t>(), nc, off_d_.as<uint32_t>(),
              adj_d_.as<uint32_t>(), src_.as<uint32_t>(), din_.as<uint32_t>(), stream_);
  // rows ascending
  if (cv->max_dout <= 8192) {
    sort_rows(off_d_.as<uint32_t>(), nc, adj_d_.as<uint32_t>(), cv->max_dout, stream_);
  } else {
    tkeys_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
    size_t tb = 0;
    KT_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, adj_d_.as<uint32_t>(),
                                                      tkeys_.as<uint32_t>(), mc, nc, off_d_.as<uint32_t>(),
                                                      off_d_.as<uint32_t>() + 1, 0, 32, stream_));
    temp_.ensure(tb);
    KT_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(temp_.p, tb, adj_d_.as<uint32_t>(),
                                                      tkeys_.as<uint32_t>(), mc, nc, off_d_.as<uint32_t>(),
                                                      off_d_.as<uint32_t>() + 1, 0, 32, stream_));
    KT_HIP(hipMemcpyAsync(adj_d_.p, tkeys_.p, sizeof(uint32_t) * mc, hipMemcpyDeviceToDevice, stream_));
  }
  // transpose, stable by destination: in-rows come out sorted by source
  tkeys_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  tvals_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  tiota_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  in_src_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  iota_u32(tiota_.as<uint32_t>(), mc, stream_);
  if (mc) {
    int bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) < nc) ++bits;
    size_t tb = 0;
    KT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, adj_d_.as<uint32_t>(), tkeys_.as<uint32_t>(),
                                              tiota_.as<uint32_t>(), tvals_.as<uint32_t>(), mc, 0, bits,
                                              stream_));
    temp_.ensure(tb);
    KT_HIP(hipcub::DeviceRadixSort::SortPairs(temp_.p, tb, adj_d_.as<uint32_t>(), tkeys_.as<uint32_t>(),
                                              tiota_.as<uint32_t>(), tvals_.as<uint32_t>(), mc, 0, bits,
                                              stream_));
  }
  gather_u32(tvals_.as<uint32_t>(), src_.as<uint32_t>(), mc, in_src_.as<uint32_t>(), stream_);
  exclusive_scan_u32(din_.as<uint32_t>(), din_off_.as<uint32_t>(), nc + 1);
  adj_u_.ensure(sizeof(uint32_t) * std::max<uint64_t>(2ull * mc, 1));
  eid_u_.ensure(sizeof(uint32_t) * std::max<uint64_t>(2ull * mc, 1));
  sup_.ensure(sizeof(int32_t) * std::max<uint32_t>(mc, 1));
  state_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  eq0_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  eq1_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  // off_u[w] = in-slots before w + out-slots before w
  add_offsets(din_off_.as<uint32_t>(), off_d_.as<uint32_t>(), nc + 1, off_u_.as<uint32_t>(), stream_);
  build_undirected(off_d_.as<uint32_t>(), adj_d_.as<uint32_t>(), din_off_.as<uint32_t>(),
                   in_src_.as<uint32_t>(), tvals_.as<uint32_t>(), nc, off_u_.as<uint32_t>(),
                   adj_u_.as<uint32_t>(), eid_u_.as<uint32_t>(), stream_);
  cv->inv = inv_.as<uint32_t>();
  cv->off_d = off_d_.as<uint32_t>();
  cv->adj_d = adj_d_.as<uint32_t>();
  cv->src = src_.as<uint32_t>();
  cv->off_u = off_u_.as<uint32_t>();
  cv->adj_u = adj_u_.as<uint32_t>();
  cv->eid_u = eid_u_.as<uint32_t>();
  cv->sup = sup_.as<int32_t>();
  cv->state = state_.as<uint32_t>();
}

// ------------------------------------------------------------ (c) count ---

void Solver::count_support(CoreView* cv, float* ms) {
  fill_u32(reinterpret_cast<uint32_t*>(cv->sup), cv->mc, 0, stream_);
  const double avg = cv->nc ? double(cv->mc) / cv->nc : 0.0;
  const int lanes = lanes_count_ ? lanes_count_ : auto_lanes(avg);
  KT_HIP(hipEventRecord(ev_[6], stream_));
  if (!force_edge_count_ && support_bitmap_fits(cv->nc, cv->max_dout))
    support_bitmap(cv->off_d, cv->adj_d, cv->nc, cv->max_dout, lanes, cv->sup, ctl_, stream_);
  else
    support_edges(cv->off_d, cv->adj_d, cv->src, cv->nc, cv->mc, lanes, cv->sup, stream_);
  KT_HIP(hipEventRecord(ev_[7], stream_));
  if (ms) {
    KT_HIP(hipEventSynchronize(ev_[7]));
    KT_HIP(hipEventElapsedTime(ms, ev_[6], ev_[7]));
  }
}

// ------------------------------------------------------------- (d) peel ---

// Truss levels on the core view, upward from threshold thr0 (an edge stays
// while its support >= thr, i.e. in the (thr+2)-truss). Empty levels are
// skipped by jumping to the minimum live support. With early_ans != 0, a
// first level of exactly early_ans edges at kmax early_k ends the search
// (src/kc.cu:1240-1241).
TrussOutcome Solver::truss_search(CoreView* cv, int32_t thr0, uint64_t early_ans, uint32_t early_k) {
  TrussOutcome out;
  const uint32_t mc = cv->mc;
  if (mc == 0) return out;
  fill_u32(cv->state, mc, kNone, stream_);
  uint32_t* qc = ctl_->qcount;
  uint32_t* q[2] = {eq0_.as<uint32_t>(), eq1_.as<uint32_t>()};
  const double avg = cv->nc ? double(2ull * mc) / cv->nc : 0.0;
  const int lanes = lanes_peel_ ? lanes_peel_ : auto_lanes(avg);
  uint64_t alive = mc;
  uint32_t r = 1;
  int32_t thr = thr0;
  bool first = true;
  for (;;) {
    zero_u32(&qc[0]);
    edge_frontier_init(cv->state, cv->sup, mc, thr, q[0], &qc[0], stream_);
    int cur = 0;
    uint32_t c = read_u32(&qc[0]);
    while (c) {
      if (c == alive) {  // every live edge goes: the level is empty
        edge_stamp(q[cur], &qc[cur], c, r, cv->state, stream_);
        ++r;
        alive = 0;
        ++out.rounds;
        break;
      }
      edge_stamp(q[cur], &qc[cur], c, r, cv->state, stream_);
      zero_u32(&qc[cur ^ 1]);
      peel_process(cv->off_u, cv->adj_u, cv->eid_u, cv->src, cv->adj_d, q[cur], &qc[cur], c, r, thr, lanes,
                   cv->state, cv->sup, q[cur ^ 1], &qc[cur ^ 1], stream_);
      alive -= c;
      ++r;
      ++out.rounds;
      cur ^= 1;
      c = read_u32(&qc[cur]);
    }
    if (alive == 0) break;
    fill_u32(&ctl_->minsup, 1, kNone, stream_);
    edge_min_support(cv->state, cv->sup, mc, &ctl_->minsup, stream_);
    const uint32_t ms = read_u32(&ctl_->minsup);
    out.kmax = ms + 2;
    out.count = alive;
    out.stamp_floor = r - 1;
    ++out.levels;
    if (first && early_ans && alive == early_ans && out.kmax == early_k) break;
    first = false;
    thr = int32_t(ms) + 1;
  }
  return out;
}

// ------------------------------------------------------------ full solve ---

void Solver::solve_pairs_device(const uint32_t* d_pairs, uint64_t m, uint32_t n, kt_result* res) {
  KT_HIP(hipSetDevice(device_));
  std::memset(res, 0, sizeof(*res));
  KT_HIP(hipEventRecord(ev_[0], stream_));
  build_from_pairs(d_pairs, m, n);
  solve_csr(res);
}

void Solver::solve_lines_device(const uint32_t* d_xs, const uint32_t* d_ys, uint64_t L, uint32_t n,
                                kt_result* res) {
  KT_HIP(hipSetDevice(device_));
  std::memset(res, 0, sizeof(*res));
  KT_HIP(hipEventRecord(ev_[0], stream_));
  build_from_lines(d_xs, d_ys, L, n);
  solve_csr(res);
}

void Solver::solve_csr(kt_result* res) {
  const uint32_t n = n_;
  res->n = n;
  res->m = m_;
  have_truss_ = false;
  KT_HIP(hipEventRecord(ev_[1], stream_));
  float t_kcore = 0, t_tmp = 0;

  // (b) vertex peel: isolated vertices first, then the galloping max-core
  lvl_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  q0_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  q1_.ensure(sizeof(uint32_t) * (size_t(n) + 1));
  degrees_from_offsets(off(), n, deg(), stream_);
  fill_u32(lvl(), n, kNone, stream_);
  const uint64_t gone = peel_vertices(1);
  if (gone >= n) {  // no edges
    KT_HIP(hipEventRecord(ev_[2], stream_));
    KT_HIP(hipEventSynchronize(ev_[2]));
    KT_HIP(hipEventElapsedTime(&res->ms_total, ev_[0], ev_[2]));
    KT_HIP(hipEventElapsedTime(&res->ms_build, ev_[0], ev_[1]));
    last_core_ = CoreView{};
    have_truss_ = true;
    return;
  }
  std::vector<uint32_t> ok{1};
  res->max_core = max_core(uint64_t(n) - gone, &ok);
  KT_HIP(hipEventRecord(ev_[2], stream_));

  // pass 1 on the max-core
  CoreView cv;
  extract_core(&cv);
  count_support(&cv, nullptr);
  TrussOutcome p1 = truss_search(&cv, 1, 0, 0);
  res->k_hat = p1.kmax;
  res->k_hat_edges = p1.count;
  KT_HIP(hipEventRecord(ev_[3], stream_));

  // pass 2 on the (L-1)-core: reopen everything removed above the largest
  // galloping threshold <= L-1, recount degrees, peel at L-1.
  const uint32_t L = std::max<uint32_t>(p1.kmax, 3);
  res->core_k = L - 1;
  uint32_t kj = 1;
  for (uint32_t t : ok)
    if (t <= L - 1) kj = std::max(kj, t);
  core_reopen_above(lvl(), n, kj, stream_);
  core_alive_degrees(off(), adj(), lvl(), n, deg(), stream_);
  if (L - 1 > kj) peel_vertices(L - 1);
  KT_HIP(hipEventRecord(ev_[4], stream_));
  extract_core(&cv);
  res->n_core = cv.nc;
  res->m_core = cv.mc;
  KT_HIP(hipEventRecord(ev_[5], stream_));
  count_support(&cv, &res->ms_count_kernel);
  {
    const double avg = cv.nc ? double(cv.mc) / cv.nc : 0.0;
    res->lanes = lanes_count_ ? lanes_count_ : auto_lanes(avg);
  }
  hipEvent_t e_count;
  KT_HIP(hipEventCreate(&e_count));
  KT_HIP(hipEventRecord(e_count, stream_));
  TrussOutcome p2 = truss_search(&cv, int32_t(L) - 2, p1.kmax >= 3 ? p1.count : 0, p1.kmax);
  hipEvent_t e_end;
  KT_HIP(hipEventCreate(&e_end));
  KT_HIP(hipEventRecord(e_end, stream_));
  KT_HIP(hipEventSynchronize(e_end));
  res->kmax = p2.kmax;
  res->truss_edges = p2.kmax ? p2.count : 0;
  res->peel_rounds = p1.rounds + p2.rounds;
  res->levels = p1.levels + p2.levels;
  KT_HIP(hipEventElapsedTime(&res->ms_total, ev_[0], e_end));
  KT_HIP(hipEventElapsedTime(&res->ms_build, ev_[0], ev_[1]));
  KT_HIP(hipEventElapsedTime(&t_kcore, ev_[1], ev_[2]));
  KT_HIP(hipEventElapsedTime(&res->ms_pass1, ev_[2], ev_[3]));
  KT_HIP(hipEventElapsedTime(&t_tmp, ev_[3], ev_[4]));
  res->ms_kcore = t_kcore + t_tmp;
  KT_HIP(hipEventElapsedTime(&res->ms_extract, ev_[4], ev_[5]));
  KT_HIP(hipEventElapsedTime(&res->ms_count, ev_[5], e_count));
  KT_HIP(hipEventElapsedTime(&res->ms_peel, e_count, e_end));
  (void)hipEventDestroy(e_count);
  (void)hipEventDestroy(e_end);
  last_core_ = cv;
  last_floor_ = p2.stamp_floor;
  last_truss_ = res->truss_edges;
  have_truss_ = true;
  if (p2.kmax == 0) last_core_.mc = 0;
}

void Solver::last_stats(kt_stats* out) {
  if (!have_truss_) fail("no solve has completed on this context");
  std::memset(out, 0, sizeof(*out));
  const CoreView& cv = last_core_;
  out->n_core = cv.nc;
  out->m_core = cv.mc;
  out->max_out_degree = cv.max_dout;
  out->truss_edges = last_truss_;
  if (cv.mc == 0) return;
  DevBuf tmp;
  tmp.ensure(16);
  fill_u64(tmp.as<uint64_t>(), 2, 0, stream_);
  w_stats(cv.off_d, cv.src, cv.adj_d, cv.state, cv.mc, last_floor_, tmp.as<uint64_t>(), stream_);
  uint64_t h[2];
  KT_HIP(hipMemcpyAsync(h, tmp.p, 16, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  tmp.release();
  out->w_core = h[0];
  out->w_peeled = h[1];
}

uint64_t Solver::truss_edges(uint32_t* host_pairs, uint64_t cap) {
  if (!have_truss_) fail("no solve has completed on this context");
  const CoreView& cv = last_core_;
  if (cv.mc == 0) return 0;
  out_pairs_.ensure(sizeof(uint32_t) * 2 * size_t(cv.mc));
  zero_u32(&ctl_->nalive);
  edges_alive_above(cv.state, cv.mc, last_floor_, cv.src, cv.adj_d, cv.inv, out_pairs_.as<uint32_t>(),
                    &ctl_->nalive, stream_);
  const uint32_t cnt = read_u32(&ctl_->nalive);
  const uint64_t take = std::min<uint64_t>(cnt, cap);
  if (take && host_pairs)
    KT_HIP(hipMemcpyAsync(host_pairs, out_pairs_.p, sizeof(uint32_t) * 2 * take, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  return cnt;
}

}  // namespace kt
