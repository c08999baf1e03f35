// kc_gpu — command-line drop-in for the reference's `kc_gpu -f <edge file>`
// (src/kc_gpu.cpp:209-347): same argument check, same exit code on bad usage
// (1), same final stdout line. Parse, solve and print go through the C ABI of
// libkc_amd.so. KC_VERBOSE=1 adds stage timings on stderr.
//
// Optional outputs after `-f <file>` (the reference writes neither):
//   -o <path>  the kmax-truss edge set, one edge per line "u\tv", u < v,
//              sorted by (u, v)
//   -t <path>  per-edge trussness, "u\tv\tt" for every edge with t >= 3,
//              sorted by (u, v); edges not listed have trussness 2
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ktruss.h"

namespace {

int usage() {
  std::fprintf(stderr, "Usage: -f [data_file_path] [-o truss_edges_path] [-t trussness_path]\n");
  return 1;
}

bool write_rows(const char* path, std::vector<uint32_t>& rows, int width) {
  const size_t count = rows.size() / width;
  std::vector<size_t> order(count);
  for (size_t i = 0; i < count; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const uint32_t* x = &rows[a * width];
    const uint32_t* y = &rows[b * width];
    return x[0] != y[0] ? x[0] < y[0] : x[1] < y[1];
  });
  FILE* f = std::fopen(path, "w");
  if (!f) return false;
  bool ok = true;
  for (size_t i : order) {
    const uint32_t* r = &rows[i * width];
    const int w = width == 2 ? std::fprintf(f, "%u\t%u\n", r[0], r[1]) : std::fprintf(f, "%u\t%u\t%u\n", r[0], r[1], r[2]);
    if (w < 0) {
      ok = false;
      break;
    }
  }
  return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 3 || argc % 2 == 0 || std::strcmp(argv[1], "-f") != 0) return usage();
  const char* out_edges = nullptr;
  const char* out_truss = nullptr;
  for (int i = 3; i < argc; i += 2) {
    if (std::strcmp(argv[i], "-o") == 0)
      out_edges = argv[i + 1];
    else if (std::strcmp(argv[i], "-t") == 0)
      out_truss = argv[i + 1];
    else
      return usage();
  }
  kt_context* ctx = nullptr;
  if (kt_create(0, &ctx) != 0) {
    std::fprintf(stderr, "kc_gpu: %s\n", kt_last_error());
    return 2;
  }
  kt_result r;
  double parse_ms = 0;
  if (kt_solve_tsv(ctx, argv[2], &r, &parse_ms) != 0) {
    std::fprintf(stderr, "kc_gpu: %s\n", kt_last_error());
    kt_destroy(ctx);
    return 255;
  }
  int rc = 0;
  if (out_edges) {  // before the trussness pass, which replaces the survivors
    uint64_t cnt = 0;
    std::vector<uint32_t> rows;
    if (kt_truss_edges(ctx, nullptr, 0, &cnt) == 0) {
      rows.resize(2 * cnt);
      if (kt_truss_edges(ctx, rows.data(), cnt, &cnt) != 0) rc = 3;
    } else {
      rc = 3;
    }
    if (rc == 0 && !write_rows(out_edges, rows, 2)) rc = 4;
  }
  if (rc == 0 && out_truss) {
    uint64_t cnt = 0;
    std::vector<uint32_t> rows;
    if (kt_trussness(ctx, nullptr, 0, &cnt) == 0) {
      rows.resize(3 * cnt);
      if (kt_trussness(ctx, rows.data(), cnt, &cnt) != 0) rc = 3;
    } else {
      rc = 3;
    }
    if (rc == 0 && !write_rows(out_truss, rows, 3)) rc = 4;
  }
  if (rc == 3) std::fprintf(stderr, "kc_gpu: %s\n", kt_last_error());
  if (rc == 4) std::fprintf(stderr, "kc_gpu: cannot write an output file\n");
  std::printf("kmax = %u, Edges in kmax-truss = %llu.\n", r.kmax, (unsigned long long)r.truss_edges);
  const char* v = std::getenv("KC_VERBOSE");
  if (v && v[0] == '1') {
    std::fprintf(stderr,
                 "n=%u m=%llu parse=%.3fms solve=%.3fms build=%.3f kcore=%.3f pass1=%.3f extract=%.3f "
                 "count=%.3f peel=%.3f max_core=%u k_hat=%u core=(%u,%llu) rounds=%llu\n",
                 r.n, (unsigned long long)r.m, parse_ms, r.ms_total, r.ms_build, r.ms_kcore, r.ms_pass1,
                 r.ms_extract, r.ms_count, r.ms_peel, r.max_core, r.k_hat, r.n_core,
                 (unsigned long long)r.m_core, (unsigned long long)r.peel_rounds);
  }
  kt_destroy(ctx);
  return rc;
}
