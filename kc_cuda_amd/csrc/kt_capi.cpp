// extern "C" surface of libkc_amd.so (declared in include/ktruss.h).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>

#include "../../include/ktruss.h"
#include "kt_batch_host.h"
#include "kt_noalloc.h"
#include "kt_serve.h"
#include "kt_solver.h"
#include "kt_tsv.h"

struct kt_context {
  std::unique_ptr<kt::Solver> solver;
  std::unique_ptr<kt::BatchRunner> batch;
  kt::BatchRunner& batcher() {
    if (!batch) batch.reset(new kt::BatchRunner(solver.get()));
    return *batch;
  }
};

namespace {
thread_local std::string g_err;

int set_err(const std::string& s, int code = 1) {
  g_err = s;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return 0;
  } catch (const kt::Error& e) {
    return set_err(e.what, int(e.code) ? int(e.code) : 1);
  } catch (const std::bad_alloc&) {
    return set_err("host allocation failed", 2);
  } catch (const std::exception& e) {
    return set_err(e.what(), 3);
  } catch (...) {
    return set_err("unknown failure", 4);
  }
}
}  // namespace

extern "C" {

const char* kt_last_error(void) { return g_err.c_str(); }

int kt_device_count(int* count) {
  return guarded([&] {
    int c = 0;
    KT_HIP(hipGetDeviceCount(&c));
    *count = c;
  });
}

int kt_create(int device, kt_context** out) {
  if (!out) return set_err("kt_create: out is NULL");
  *out = nullptr;
  return guarded([&] {
    auto* c = new kt_context;
    try {
      c->solver = std::make_unique<kt::Solver>(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int kt_destroy(kt_context* ctx) {
  return guarded([&] { delete ctx; });
}

int kt_malloc(kt_context* ctx, uint64_t bytes, void** dptr) {
  return guarded([&] {
    KT_HIP(hipSetDevice(ctx->solver->device()));
    KT_HIP(hipMalloc(dptr, bytes ? bytes : 1));
  });
}

int kt_free(kt_context* ctx, void* dptr) {
  return guarded([&] {
    KT_HIP(hipSetDevice(ctx->solver->device()));
    KT_HIP(hipFree(dptr));
  });
}

int kt_memcpy_h2d(kt_context* ctx, void* dst, const void* src, uint64_t bytes) {
  return guarded([&] {
    KT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->solver->stream()));
    KT_HIP(hipStreamSynchronize(ctx->solver->stream()));
  });
}

int kt_memcpy_d2h(kt_context* ctx, void* dst, const void* src, uint64_t bytes) {
  return guarded([&] {
    KT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->solver->stream()));
    KT_HIP(hipStreamSynchronize(ctx->solver->stream()));
  });
}

int kt_memcpy_d2d(kt_context* ctx, void* dst, const void* src, uint64_t bytes) {
  return guarded([&] {
    KT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->solver->stream()));
    KT_HIP(hipStreamSynchronize(ctx->solver->stream()));
  });
}

int kt_synchronize(kt_context* ctx) {
  return guarded([&] { KT_HIP(hipStreamSynchronize(ctx->solver->stream())); });
}

int kt_solve_pairs_device(kt_context* ctx, const uint32_t* d_pairs, uint64_t m, uint32_t n,
                          kt_result* res) {
  if (!ctx || !res) return set_err("kt_solve_pairs_device: NULL argument");
  return guarded([&] { ctx->solver->solve_pairs_device(d_pairs, m, n, res); });
}

int kt_solve_pairs_host(kt_context* ctx, const uint32_t* pairs, uint64_t m, uint32_t n,
                        kt_result* res) {
  if (!ctx || !res) return set_err("kt_solve_pairs_host: NULL argument");
  return guarded([&] {
    auto& s = *ctx->solver;
    KT_HIP(hipSetDevice(s.device()));
    void* d = nullptr;
    KT_HIP(hipMalloc(&d, sizeof(uint32_t) * 2 * (m ? m : 1)));
    try {
      if (m) KT_HIP(hipMemcpyAsync(d, pairs, sizeof(uint32_t) * 2 * m, hipMemcpyHostToDevice, s.stream()));
      s.solve_pairs_device(static_cast<uint32_t*>(d), m, n, res);
    } catch (...) {
      (void)hipStreamSynchronize(s.stream());
      (void)hipFree(d);
      throw;
    }
    KT_HIP(hipStreamSynchronize(s.stream()));
    KT_HIP(hipFree(d));
  });
}

int kt_solve_lines_host(kt_context* ctx, const uint32_t* xs, const uint32_t* ys, uint64_t L,
                        uint32_t n, kt_result* res) {
  if (!ctx || !res) return set_err("kt_solve_lines_host: NULL argument");
  return guarded([&] {
    auto& s = *ctx->solver;
    KT_HIP(hipSetDevice(s.device()));
    void* d = nullptr;
    KT_HIP(hipMalloc(&d, sizeof(uint32_t) * 2 * (L ? L : 1)));
    uint32_t* dx = static_cast<uint32_t*>(d);
    uint32_t* dy = dx + (L ? L : 1);
    try {
      if (L) {
        KT_HIP(hipMemcpyAsync(dx, xs, sizeof(uint32_t) * L, hipMemcpyHostToDevice, s.stream()));
        KT_HIP(hipMemcpyAsync(dy, ys, sizeof(uint32_t) * L, hipMemcpyHostToDevice, s.stream()));
      }
      s.solve_lines_device(dx, dy, L, n, res);
    } catch (...) {
      (void)hipStreamSynchronize(s.stream());
      (void)hipFree(d);
      throw;
    }
    KT_HIP(hipStreamSynchronize(s.stream()));
    KT_HIP(hipFree(d));
  });
}

int kt_parse_tsv(const char* path, uint32_t* xs, uint32_t* ys, uint64_t cap, uint64_t* lines,
                 uint32_t* n) {
  return guarded([&] {
    kt::ParsedLines p = kt::parse_tsv_file(path);
    if (lines) *lines = p.xs.size();
    if (n) *n = p.n;
    if (xs && ys) {
      const uint64_t take = std::min<uint64_t>(cap, p.xs.size());
      std::memcpy(xs, p.xs.data(), sizeof(uint32_t) * take);
      std::memcpy(ys, p.ys.data(), sizeof(uint32_t) * take);
    }
  });
}

int kt_parse_tsv_device(kt_context* ctx, const char* path, uint32_t* xs, uint32_t* ys, uint64_t cap,
                        uint64_t* lines, uint32_t* n, int* on_device) {
  if (!ctx || !path) return set_err("kt_parse_tsv_device: NULL argument");
  return guarded([&] {
    uint64_t L = 0;
    uint32_t nn = 0;
    const bool dev = ctx->solver->ingest_tsv(path, &L, &nn);
    if (on_device) *on_device = dev ? 1 : 0;
    if (!dev) {  // a line with a single number: the host parser's reading
      kt::ParsedLines p = kt::parse_tsv_file(path);
      L = p.xs.size();
      nn = p.n;
      if (xs && ys) {
        const uint64_t take = std::min<uint64_t>(cap, L);
        std::memcpy(xs, p.xs.data(), sizeof(uint32_t) * take);
        std::memcpy(ys, p.ys.data(), sizeof(uint32_t) * take);
      }
    } else if (xs && ys) {
      ctx->solver->copy_ingested(xs, ys, std::min<uint64_t>(cap, L));
    }
    if (lines) *lines = L;
    if (n) *n = nn;
  });
}

int kt_solve_tsv(kt_context* ctx, const char* path, kt_result* res, double* parse_ms) {
  if (!ctx || !res || !path) return set_err("kt_solve_tsv: NULL argument");
  bool done = false;
  int rc = guarded([&] { done = ctx->solver->solve_tsv_device(path, res, parse_ms); });
  if (rc || done) return rc;
  kt::ParsedLines p;
  rc = guarded([&] {
    auto t0 = std::chrono::steady_clock::now();
    p = kt::parse_tsv_file(path);
    auto t1 = std::chrono::steady_clock::now();
    if (parse_ms) *parse_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  });
  if (rc) return rc;
  return kt_solve_lines_host(ctx, p.xs.data(), p.ys.data(), p.xs.size(), p.n, res);
}

int kt_truss_edges(kt_context* ctx, uint32_t* pairs, uint64_t cap, uint64_t* count) {
  if (!ctx) return set_err("kt_truss_edges: NULL context");
  return guarded([&] {
    uint64_t c = ctx->solver->truss_edges(pairs, cap);
    if (count) *count = c;
  });
}

int kt_push_probe(kt_context* ctx, int on) {
  if (!ctx) return set_err("kt_push_probe: NULL context");
  return guarded([&] { ctx->solver->set_push_probe(on); });
}

int kt_kernel_probe(kt_context* ctx, int kernel, int mode) {
  if (!ctx) return set_err("kt_kernel_probe: NULL context");
  return guarded([&] { ctx->solver->set_probe(kernel, mode); });
}

int kt_push_stats_get(kt_context* ctx, kt_push_stats* out) {
  if (!ctx || !out) return set_err("kt_push_stats_get: NULL argument");
  return guarded([&] { ctx->solver->push_stats(&out->launches, &out->ms, &out->slots, &out->vertices); });
}

int kt_trussness(kt_context* ctx, uint32_t* triples, uint64_t cap, uint64_t* count) {
  if (!ctx) return set_err("kt_trussness: NULL context");
  return guarded([&] {
    const uint64_t c = ctx->solver->trussness(triples, cap);
    if (count) *count = c;
  });
}

int kt_set_tuning(kt_context* ctx, int count_lanes, int peel_lanes, int flags) {
  if (!ctx) return set_err("kt_set_tuning: NULL context");
  auto ok = [](int l) { return l == 0 || l == 8 || l == 16 || l == 32 || l == 64; };
  if (!ok(count_lanes) || !ok(peel_lanes)) return set_err("kt_set_tuning: lanes must be 0, 8, 16, 32 or 64");
  ctx->solver->set_lanes(count_lanes, peel_lanes);
  ctx->solver->set_force_edge_count((flags & KT_FORCE_EDGE_COUNT) != 0);
  ctx->solver->set_force_atomic_build((flags & KT_FORCE_ATOMIC_BUILD) != 0);
  ctx->solver->set_sorted_build((flags & KT_FORCE_UNSORTED_BUILD) == 0);
  ctx->solver->set_core_mode((flags & KT_FORCE_SPARSE) ? 1 : (flags & KT_FORCE_DENSE) ? 2 : 0);
  ctx->solver->set_gallop((flags & KT_FORCE_LEVEL_WALK) == 0, (flags & KT_FORCE_PROBES) ? 1u : 16u);
  ctx->solver->set_vertex_mode((flags & KT_FORCE_VERTEX_ROUNDS) ? 1 : (flags & KT_FORCE_VERTEX_CASCADE) ? 2 : 0);
  ctx->solver->reset_drop_ratio();
  g_err.clear();
  return 0;
}

int kt_batch_solve(kt_context* ctx, const uint32_t* d_pairs, const uint64_t* h_offsets, const uint32_t* h_ns,
                   uint32_t count, uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status, uint8_t* d_keep,
                   int flags) {
  if (!ctx || (count && (!d_pairs || !h_offsets || !h_ns || !d_kmax || !d_count || !d_status)))
    return set_err("kt_batch_solve: NULL argument");
  if (flags != 0) return set_err("kt_batch_solve: flags is reserved and must be 0");
  return guarded([&] { ctx->batcher().solve(d_pairs, h_offsets, h_ns, count, d_kmax, d_count, d_status, d_keep); });
}

int kt_batch_autotune(kt_context* ctx, const uint32_t* d_pairs, const uint64_t* h_offsets, const uint32_t* h_ns,
                      uint32_t count, uint32_t sample_every, uint32_t* d_kmax, uint64_t* d_count, uint32_t* d_status,
                      kt_batch_tuning* out) {
  if (!ctx || !out || (count && (!d_pairs || !h_offsets || !h_ns || !d_kmax || !d_count || !d_status)))
    return set_err("kt_batch_autotune: NULL argument");
  return guarded([&] {
    const kt::BatchTuning t =
        ctx->batcher().autotune(d_pairs, h_offsets, h_ns, count, sample_every, d_kmax, d_count, d_status);
    out->count_lanes = t.best.count_lanes;
    out->peel_lanes = t.best.peel_lanes;
    out->edge_count = t.best.edge_count ? 1 : 0;
    out->trials = t.trials;
    out->drop_ratio = t.best.drop_ratio;
    out->ms_default = t.ms_default;
    out->ms_best = t.ms_best;
  });
}

struct kt_server {
  std::unique_ptr<kt::Server> srv;
};

int kt_serve_create(int device, int slots, uint64_t hbm_budget, kt_server** out) {
  if (!out) return set_err("kt_serve_create: NULL out");
  *out = nullptr;
  return guarded([&] {
    int count = 0;
    KT_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw kt::Error{hipErrorInvalidValue, "kt_serve_create: no such device"};
    auto s = std::make_unique<kt_server>();
    s->srv = std::make_unique<kt::Server>(device, slots, hbm_budget);
    *out = s.release();
  });
}

int kt_serve_destroy(kt_server* srv) {
  if (!srv) return 0;
  return guarded([&] { delete srv; });
}

int kt_serve_submit(kt_server* srv, const uint32_t* pairs, uint64_t m, uint32_t n, int on_device,
                    uint64_t* ticket) {
  if (!srv || !ticket) return set_err("kt_serve_submit: NULL argument");
  return guarded([&] { *ticket = srv->srv->submit(pairs, m, n, on_device != 0); });
}

int kt_serve_wait(kt_server* srv, uint64_t ticket, kt_serve_result* out) {
  if (!srv || !out) return set_err("kt_serve_wait: NULL argument");
  std::string query_err;
  const int rc = guarded([&] {
    kt::ServeOutcome o = srv->srv->wait(ticket);
    query_err = o.error;
    out->status = o.status;
    out->slot = o.slot;
    out->result = o.res;
    out->footprint = o.footprint;
    out->charged = o.charged;
    out->slot_bytes = o.slot_bytes;
    out->queued_ms = o.queued_ms;
    out->run_ms = o.run_ms;
    out->total_ms = o.total_ms;
    out->group = o.group;
  });
  if (rc == 0 && !query_err.empty()) g_err = query_err;
  return rc;
}

int kt_serve_counters_get(kt_server* srv, kt_serve_counters* out) {
  if (!srv || !out) return set_err("kt_serve_counters_get: NULL argument");
  return guarded([&] {
    const kt::ServeCounters c = srv->srv->counters();
    out->submitted = c.submitted;
    out->completed = c.completed;
    out->rejected = c.rejected;
    out->failed = c.failed;
    out->in_flight = c.in_flight;
    out->peak_in_flight = c.peak_in_flight;
    out->bytes_in_use = c.bytes_in_use;
    out->peak_bytes_in_use = c.peak_bytes_in_use;
    out->over_estimate = c.over_estimate;
    out->trims = c.trims;
    out->grouped = c.grouped;
  });
}

uint64_t kt_serve_footprint(uint64_t m, uint32_t n) { return kt::Server::footprint(m, n); }

int kt_last_stats(kt_context* ctx, kt_stats* out) {
  if (!ctx || !out) return set_err("kt_last_stats: NULL argument");
  return guarded([&] { ctx->solver->last_stats(out); });
}

int kt_last_csr(kt_context* ctx, uint64_t* off, uint32_t* adj) {
  if (!ctx) return set_err("kt_last_csr: NULL context");
  return guarded([&] { ctx->solver->csr_to_host(off, adj); });
}

int kt_solve_pairs_device_noalloc(int device, const uint32_t* d_pairs, uint64_t m, uint32_t n,
                                  uint32_t* d_truss, uint64_t cap, uint64_t* truss_count,
                                  kt_result* res) {
  if (!res || (m && !d_pairs)) return set_err("kt_solve_pairs_device_noalloc: NULL argument");
  return guarded([&] {
    const uint64_t c = kt::noalloc_solve_pairs(device, d_pairs, m, n, d_truss, cap, res);
    if (truss_count) *truss_count = c;
  });
}

}  // extern "C"
