// Reference-compatible stage functions: the eight entry points of the
// reference's src/kc_gpu.h (C++ linkage, same signatures and semantics) and
// their C-ABI twins (include/kc_gpu_c.h), implemented on the device stages of
// this library. Like the reference (src/kc.cu:584-651 and friends) every call
// takes host arrays, runs on device 0 and returns host arrays; device state
// lives in one process-wide solver, so calls must not overlap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kc_gpu.h"
#include "../../include/kc_gpu_c.h"
#include "kt_kernels.h"
#include "kt_solver.h"

namespace kt {

namespace {
std::mutex g_mu;
std::unique_ptr<Solver> g_solver;
thread_local std::string g_kc_err;

Solver& solver() {
  if (!g_solver) g_solver = std::make_unique<Solver>(0);
  return *g_solver;
}

template <typename F>
int kc_guard(F&& f) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_kc_err = e.what;
    return int(e.code) ? int(e.code) : 1;
  } catch (const std::exception& e) {
    g_kc_err = e.what();
    return 3;
  }
}

void load_csr_u32(uint32_t N, const uint32_t* nbr_u, const uint32_t* nbr) {
  std::vector<uint64_t> off(size_t(N) + 1);
  for (uint32_t i = 0; i <= N; ++i) off[i] = nbr_u[i];
  solver().load_csr_host(N, off.data(), nbr);
}

// The reference reports errors on stderr and exits with the HIP/CUDA code
// (src/kc.cu:17-23); the C++ shims keep that convention.
void exit_on(int rc) {
  if (rc) {
    std::fprintf(stderr, "GPUassert: %s\n", g_kc_err.c_str());
    std::exit(rc);
  }
}
}  // namespace

// ------------------------------------------------- solver-side helpers ---

void Solver::vertex_state_from_host(const uint32_t* deg_origin, const uint32_t* mask) {
  LevelGraph& g = level0();
  std::vector<uint32_t> l(g.n);
  bool any = false;
  for (uint32_t v = 0; v < g.n; ++v) {
    l[v] = mask && mask[v] ? 0u : kNone;
    any |= (l[v] == 0u);
  }
  KT_HIP(hipMemcpyAsync(g.lvl, l.data(), sizeof(uint32_t) * g.n, hipMemcpyHostToDevice, stream_));
  if (any || !deg_origin) {
    core_alive_degrees(g.off, g.adj, g.lvl, g.n, g.deg, heavy(), stream_);
  } else {
    KT_HIP(hipMemcpyAsync(g.deg, deg_origin, sizeof(uint32_t) * g.n, hipMemcpyHostToDevice, stream_));
  }
  KT_HIP(hipStreamSynchronize(stream_));
}

void Solver::level0_mask_to_host(uint32_t* mask) {
  LevelGraph& g = *levels_[0];
  std::vector<uint32_t> l(g.n);
  KT_HIP(hipMemcpyAsync(l.data(), g.lvl, sizeof(uint32_t) * g.n, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  for (uint32_t v = 0; v < g.n; ++v) mask[v] = (l[v] != kNone) ? 1u : 0u;
}

void Solver::level0_degrees_to_host(uint32_t* deg) {
  LevelGraph& g = *levels_[0];
  KT_HIP(hipMemcpyAsync(deg, g.deg, sizeof(uint32_t) * g.n, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
}

void Solver::core_ranks_to_host(const CoreView& cv, uint32_t* inv_orig, uint32_t* dout) {
  if (cv.nc) {
    KT_HIP(hipMemcpyAsync(inv_orig, cv.inv, sizeof(uint32_t) * cv.nc, hipMemcpyDeviceToHost, stream_));
    std::vector<uint32_t> off(size_t(cv.nc) + 1);
    KT_HIP(hipMemcpyAsync(off.data(), cv.off_d0, sizeof(uint32_t) * (cv.nc + 1), hipMemcpyDeviceToHost, stream_));
    KT_HIP(hipStreamSynchronize(stream_));
    for (uint32_t r = 0; r < cv.nc; ++r) dout[r] = off[r + 1] - off[r];
  }
}

// Oriented rows for caller-given ranks `to` of the kept vertices (the
// reference's build_nbr_direct, src/kc.cu:1478-1580): row to[u] holds
// {to[v] : v kept, to[v] > to[u]}, ascending, at the caller's offsets.
void Solver::orient_with_ranks(const uint32_t* h_to, uint32_t new_n, const uint32_t* h_off_d, uint32_t* h_adj_d) {
  core_valid_ = false;  // the core buffers are rewritten
  LevelGraph& g = *levels_[0];
  std::vector<uint32_t> l(g.n), to(g.n), inv(size_t(new_n) + 1, 0);
  KT_HIP(hipMemcpyAsync(l.data(), g.lvl, sizeof(uint32_t) * g.n, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  for (uint32_t v = 0; v < g.n; ++v) {
    to[v] = (l[v] == kNone) ? h_to[v] : kNone;
    if (l[v] == kNone) {
      if (h_to[v] >= new_n) fail("build_nbr_direct: rank out of range");
      inv[h_to[v]] = v;
    }
  }
  const uint32_t mc = h_off_d[new_n];
  to_.ensure(sizeof(uint32_t) * (size_t(g.n) + 1));
  inv_.ensure(sizeof(uint32_t) * (size_t(new_n) + 1));
  off_d_.ensure(sizeof(uint32_t) * (size_t(new_n) + 1));
  adj_d_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  src_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  dout_.ensure(sizeof(uint32_t) * (size_t(new_n) + 1));
  KT_HIP(hipMemcpyAsync(to_.p, to.data(), sizeof(uint32_t) * g.n, hipMemcpyHostToDevice, stream_));
  KT_HIP(hipMemcpyAsync(inv_.p, inv.data(), sizeof(uint32_t) * new_n, hipMemcpyHostToDevice, stream_));
  KT_HIP(hipMemcpyAsync(off_d_.p, h_off_d, sizeof(uint32_t) * (size_t(new_n) + 1), hipMemcpyHostToDevice, stream_));
  fill_u32(dout_.as<uint32_t>(), uint64_t(new_n) + 1, 0, stream_);
  orient_count(g.off, g.adj, to_.as<uint32_t>(), inv_.as<uint32_t>(), new_n, dout_.as<uint32_t>(), stream_);
  std::vector<uint32_t> d(size_t(new_n) + 1);
  KT_HIP(hipMemcpyAsync(d.data(), dout_.p, sizeof(uint32_t) * new_n, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  uint32_t mx = 0;
  for (uint32_t r = 0; r < new_n; ++r) {
    if (h_off_d[r + 1] - h_off_d[r] != d[r]) fail("build_nbr_direct: offsets do not match the ranks");
    mx = std::max(mx, d[r]);
  }
  orient_fill(g.off, g.adj, to_.as<uint32_t>(), inv_.as<uint32_t>(), new_n, off_d_.as<uint32_t>(),
              adj_d_.as<uint32_t>(), src_.as<uint32_t>(), nullptr, stream_);
  sort_oriented_rows(new_n, mc, mx);
  if (mc) KT_HIP(hipMemcpyAsync(h_adj_d, adj_d_.p, sizeof(uint32_t) * mc, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
}

// Core view over a host oriented CSR in rank ids (rows ascending, u < v).
void Solver::view_from_oriented_host(uint32_t nc, const uint32_t* h_off_d, const uint32_t* h_adj_d, CoreView* cv) {
  core_valid_ = false;  // the core buffers are rewritten
  const uint32_t mc = h_off_d[nc];
  if (mc >= 0x7FFFFFFFu) fail("oriented graph exceeds 2^31 edges");
  uint32_t mx = 0;
  for (uint32_t r = 0; r < nc; ++r) mx = std::max(mx, h_off_d[r + 1] - h_off_d[r]);
  const size_t mm = std::max<uint32_t>(mc, 1);
  off_d_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  adj_d_.ensure(sizeof(uint32_t) * mm);
  src_.ensure(sizeof(uint32_t) * mm);
  inv_orig_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  orig0_.ensure(sizeof(uint32_t) * mm);
  del_thr_.ensure(sizeof(uint32_t) * mm);
  sup_.ensure(sizeof(int32_t) * mm);
  state_.ensure(sizeof(uint32_t) * mm);
  eq0_.ensure(sizeof(uint32_t) * mm);
  eq1_.ensure(sizeof(uint32_t) * mm);
  KT_HIP(hipMemcpyAsync(off_d_.p, h_off_d, sizeof(uint32_t) * (size_t(nc) + 1), hipMemcpyHostToDevice, stream_));
  if (mc) KT_HIP(hipMemcpyAsync(adj_d_.p, h_adj_d, sizeof(uint32_t) * mc, hipMemcpyHostToDevice, stream_));
  row_src(off_d_.as<uint32_t>(), nc, src_.as<uint32_t>(), stream_);
  iota_u32(inv_orig_.as<uint32_t>(), nc, stream_);
  iota_u32(orig0_.as<uint32_t>(), mc, stream_);
  fill_u32(del_thr_.as<uint32_t>(), mc, kNone, stream_);
  fill_u32(state_.as<uint32_t>(), mc, kNone, stream_);
  *cv = CoreView{};
  cv->nc = cv->nc0 = nc;
  cv->mc = cv->mc0 = mc;
  cv->max_dout = cv->max_dout0 = mx;
  cv->inv = inv_orig_.as<uint32_t>();
  cv->off_d = cv->off_d0 = off_d_.as<uint32_t>();
  cv->adj_d = cv->adj_d0 = adj_d_.as<uint32_t>();
  cv->src = cv->src0 = src_.as<uint32_t>();
  cv->orig = orig0_.as<uint32_t>();
  cv->del_thr = del_thr_.as<uint32_t>();
  cv->sup = sup_.as<int32_t>();
  cv->state = state_.as<uint32_t>();
  cv->work = -1;
  build_undirected_view(cv);
  if (core_mode_ == 2 ? nc <= 65536 : (core_mode_ == 0 && dense_fits(nc, mc))) make_dense(cv);
}

void Solver::undirected_to_host(const CoreView& cv, uint32_t* adj_u, uint32_t* eid_u, uint32_t* lson,
                                uint32_t* rson) {
  const size_t mm = std::max<uint32_t>(cv.mc, 1);
  DevBuf ls, rs;
  ls.ensure(sizeof(uint32_t) * mm);
  rs.ensure(sizeof(uint32_t) * mm);
  sons(cv.off_u, cv.adj_u, cv.eid_u, cv.nc, ls.as<uint32_t>(), rs.as<uint32_t>(), stream_);
  const size_t m2 = size_t(cv.mc) * 2;
  if (m2) {
    if (adj_u) KT_HIP(hipMemcpyAsync(adj_u, cv.adj_u, sizeof(uint32_t) * m2, hipMemcpyDeviceToHost, stream_));
    if (eid_u) KT_HIP(hipMemcpyAsync(eid_u, cv.eid_u, sizeof(uint32_t) * m2, hipMemcpyDeviceToHost, stream_));
    if (lson) KT_HIP(hipMemcpyAsync(lson, ls.p, sizeof(uint32_t) * cv.mc, hipMemcpyDeviceToHost, stream_));
    if (rson) KT_HIP(hipMemcpyAsync(rson, rs.p, sizeof(uint32_t) * cv.mc, hipMemcpyDeviceToHost, stream_));
  }
  KT_HIP(hipStreamSynchronize(stream_));
}

void Solver::support_to_host(const CoreView& cv, int32_t* sup) {
  if (cv.mc) KT_HIP(hipMemcpyAsync(sup, cv.sup, sizeof(int32_t) * cv.mc, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
}

}  // namespace kt

using kt::solver;

// ------------------------------------------------------------- C ABI ---

extern "C" {

const char* kc_last_error(void) { return kt::g_kc_err.c_str(); }

int kc_initGPU(uint32_t edge_num, uint32_t N) {
  return kt::kc_guard([&] {
    kt::Solver& s = solver();
    const char* v = std::getenv("KC_VERBOSE");
    if (v && v[0] == '1') {
      size_t fr = 0, tot = 0;
      KT_HIP(hipMemGetInfo(&fr, &tot));
      std::fprintf(stderr, "device %d: %zu MiB total, %zu MiB free; graph needs about %llu MiB\n", s.device(),
                   tot >> 20, fr >> 20,
                   (unsigned long long)((uint64_t(edge_num) * 40 + uint64_t(N + 1) * 16) >> 20));
    }
  });
}

int kc_kmax_core_gpu(uint32_t N, const uint32_t* nbr_u, const uint32_t* nbr, const uint32_t* deg_origin,
                     uint32_t* mask, uint32_t launch_start, uint32_t* result) {
  return kt::kc_guard([&] {
    kt::load_csr_u32(N, nbr_u, nbr);
    kt::Solver& s = solver();
    s.vertex_state_from_host(deg_origin, mask);
    if (launch_start == 0xFFFFFFFFu) {
      // max-core; returns its k minus one (src/kc.cu:870-883)
      const uint32_t c = s.max_core(false, false);
      s.level0_mask_to_host(mask);
      *result = c ? c - 1 : 0;
    } else if (launch_start == 0) {
      *result = 0;
    } else {
      s.peel_vertices(s.level0(), launch_start);
      s.level0_mask_to_host(mask);
      *result = launch_start - 1;
    }
  });
}

int kc_count_deg_with_mask(const uint32_t* nbr, const uint32_t* nbr_u, uint32_t N, const uint32_t* node_delete,
                           uint32_t* deg_with_mask, uint32_t* to, kc_link* li, uint32_t* new_N, uint32_t* deg,
                           uint32_t* deg_direct) {
  return kt::kc_guard([&] {
    kt::load_csr_u32(N, nbr_u, nbr);
    kt::Solver& s = solver();
    s.vertex_state_from_host(nullptr, node_delete);
    std::vector<uint32_t> dm(N);
    s.level0_degrees_to_host(dm.data());
    kt::CoreView cv;
    s.extract_core(s.level0(), &cv);
    const uint32_t nc = cv.nc;
    std::vector<uint32_t> inv(nc), dout(nc), pos(N, 0);
    s.core_ranks_to_host(cv, inv.data(), dout.data());
    uint32_t k = 0;
    for (uint32_t v = 0; v < N; ++v)
      if (!node_delete[v]) {
        pos[v] = k++;
        deg_with_mask[v] += dm[v];
      }
    for (uint32_t r = 0; r < nc; ++r) {
      const uint32_t v = inv[r];
      li[r].x = v;
      li[r].t = pos[v];
      li[r].d = deg_with_mask[v];
      to[v] = r;
      deg[r] = deg_with_mask[v];
      deg_direct[r] = dout[r];
    }
    *new_N = nc;
  });
}

int kc_build_nbr_direct(const uint32_t* nbr, const uint32_t* nbr_u, uint32_t N, const uint32_t* node_delete,
                        const uint32_t* to, const uint32_t* nbr_u_direct, uint32_t* nbr_direct, uint32_t new_N) {
  return kt::kc_guard([&] {
    kt::load_csr_u32(N, nbr_u, nbr);
    kt::Solver& s = solver();
    s.vertex_state_from_host(nullptr, node_delete);
    s.orient_with_ranks(to, new_N, nbr_u_direct, nbr_direct);
  });
}

int kc_build_graph_from_direct(const uint32_t* nbr_direct, const uint32_t* nbr_u_direct, uint32_t N, uint32_t* nbr,
                               uint32_t* fa, uint32_t* left_son, uint32_t* right_son) {
  return kt::kc_guard([&] {
    kt::Solver& s = solver();
    kt::CoreView cv;
    s.view_from_oriented_host(N, nbr_u_direct, nbr_direct, &cv);
    s.undirected_to_host(cv, nbr, fa, left_son, right_son);
  });
}

int kc_tricount_direct(const uint32_t* nbr_u_direct, const uint32_t* nbr_direct, uint32_t N, uint32_t* nbr_tc,
                       const uint32_t* left_son, const uint32_t* right_son, uint32_t* nbr_tc_double,
                       uint32_t tc_limit) {
  return kt::kc_guard([&] {
    kt::Solver& s = solver();
    kt::CoreView cv;
    s.view_from_oriented_host(N, nbr_u_direct, nbr_direct, &cv);
    s.count_support(&cv, nullptr);
    std::vector<int32_t> sup(cv.mc);
    s.support_to_host(cv, sup.data());
    // Exact supports, doubled. With tc_limit != 0 the reference returns
    // pruned lower bounds (src/kc.cu:537-582) that are only ever used to drop
    // edges below tc_limit; exact values satisfy that contract too.
    for (uint32_t e = 0; e < cv.mc; ++e) nbr_tc[e] = 2u * uint32_t(sup[e]);
    if (tc_limit == 0 && nbr_tc_double && left_son && right_son)
      for (uint32_t e = 0; e < cv.mc; ++e) nbr_tc_double[left_son[e]] = nbr_tc_double[right_son[e]] = nbr_tc[e];
  });
}

int kc_kmax_truss_GPU(uint32_t* k_max, uint32_t* k_max_num, uint32_t ans, uint32_t N, uint32_t k, uint32_t k_step,
                      const uint32_t* nbr_u_direct, const uint32_t* nbr_direct) {
  return kt::kc_guard([&] {
    kt::Solver& s = solver();
    // Doubled thresholds: the reference peels levels k + 2*step, ... and
    // backs off towards k + 2 when a level empties (src/kc.cu:1169-1305),
    // so the reachable levels are k + 2, k + 4, ...; its first level
    // k_init = k + 2*k_step ends the search when it keeps exactly `ans`
    // edges (src/kc.cu:1240-1241).
    const uint32_t k_init = k + 2 * k_step;
    const int32_t thr_init = int32_t(k_init / 2), thr_low = int32_t(k / 2 + 1);
    kt::CoreView cv;
    s.view_from_oriented_host(N, nbr_u_direct, nbr_direct, &cv);
    s.count_support(&cv, nullptr);
    kt::TrussOutcome o = s.truss_search(&cv, thr_init, ans, k_init / 2 + 2, true);
    if (o.kmax == 0 && thr_low < thr_init) {
      s.view_from_oriented_host(N, nbr_u_direct, nbr_direct, &cv);
      s.count_support(&cv, nullptr);
      o = s.truss_search(&cv, thr_low, 0, 0, false);
    }
    if (o.kmax) {
      *k_max = o.kmax;
      *k_max_num = uint32_t(o.count);
    }
  });
}

void kc_test_cuda(void) {}

}  // extern "C"

// ---------------------------------------------------------- C++ ABI ---
// Same signatures as the reference's src/kc_gpu.h:45-69.

void initGPU(const Bint_t edge_num, const uint32_t N) { kt::exit_on(kc_initGPU(edge_num, N)); }

uint32_t kmax_core_gpu(const uint32_t N, const Bint_t* nbr_u, const uint32_t* nbr, const uint32_t* deg_origin,
                       uint32_t*& mask, uint32_t Launch_start) {
  uint32_t r = 0;
  kt::exit_on(kc_kmax_core_gpu(N, nbr_u, nbr, deg_origin, mask, Launch_start, &r));
  return r;
}

void count_deg_with_mask(const uint32_t*& nbr, const Bint_t*& nbr_u, uint32_t N, uint32_t*& node_delete,
                         uint32_t*& deg_with_mask, uint32_t*& to, Link_t*& li, uint32_t& new_N, uint32_t*& deg,
                         uint32_t*& deg_direct) {
  // deg and deg_direct are allocated here and freed by the caller
  // (src/kc.cu:1461-1462, src/kc_gpu.cpp:119-120).
  uint32_t kept = 0;
  for (uint32_t v = 0; v < N; ++v) kept += node_delete[v] == 0;
  deg = new uint32_t[kept ? kept : 1]();
  deg_direct = new uint32_t[kept ? kept : 1]();
  kt::exit_on(kc_count_deg_with_mask(nbr, nbr_u, N, node_delete, deg_with_mask, to,
                                     reinterpret_cast<kc_link*>(li), &new_N, deg, deg_direct));
}

void build_nbr_direct(const uint32_t*& nbr, const Bint_t*& nbr_u, uint32_t N, uint32_t*& node_delete, uint32_t*& to,
                      uint32_t*& deg, Bint_t*& nbr_u_direct, uint32_t*& nbr_direct, uint32_t new_N) {
  (void)deg;
  kt::exit_on(kc_build_nbr_direct(nbr, nbr_u, N, node_delete, to, nbr_u_direct, nbr_direct, new_N));
}

void build_graph_from_direct(uint32_t*& nbr_direct, Bint_t*& nbr_u_direct, uint32_t N, uint32_t*& nbr, Bint_t*& nbr_u,
                             Bint_t*& fa, Bint_t*& left_son, Bint_t*& right_son) {
  (void)nbr_u;
  kt::exit_on(kc_build_graph_from_direct(nbr_direct, nbr_u_direct, N, nbr, fa, left_son, right_son));
}

void tricount_direct(const Bint_t* nbr_u_direct, const uint32_t* nbr_direct, const uint32_t N, const Bint_t* nbr_u,
                     const uint32_t* nbr, uint32_t* nbr_tc, const Bint_t* left_son, const Bint_t* right_son,
                     uint32_t* nbr_tc_double, uint32_t tc_limit) {
  (void)nbr_u;
  (void)nbr;
  kt::exit_on(kc_tricount_direct(nbr_u_direct, nbr_direct, N, nbr_tc, left_son, right_son, nbr_tc_double, tc_limit));
}

void kmax_truss_GPU(uint32_t& k_max, Bint_t& k_max_num, Bint_t ans, uint32_t N, uint32_t k, uint32_t k_step,
                    Bint_t*& nbr_u_direct, uint32_t*& nbr_direct, Bint_t*& nbr_u, uint32_t*& nbr, Bint_t*& fa,
                    Bint_t*& left_son, Bint_t*& right_son, uint32_t*& nbr_mask, uint32_t*& nbr_mask_double,
                    uint32_t*& nbr_tc, uint32_t*& nbr_tc_double, uint32_t& accurate_tc) {
  // The undirected view, masks and counts the caller passes are rebuilt on
  // the device from the oriented CSR; the caller's arrays are left as they
  // are (it only frees them afterwards, src/kc_gpu.cpp:198-200).
  (void)nbr_u, (void)nbr, (void)fa, (void)left_son, (void)right_son, (void)nbr_mask, (void)nbr_mask_double;
  (void)nbr_tc, (void)nbr_tc_double;
  accurate_tc = 1;
  kt::exit_on(kc_kmax_truss_GPU(&k_max, &k_max_num, ans, N, k, k_step, nbr_u_direct, nbr_direct));
}

void test_cuda() { kc_test_cuda(); }
