// kc_gpu — command-line drop-in for the reference's `kc_gpu -f <edge file>`
// (src/kc_gpu.cpp:209-347): same argument check, same exit code on bad usage
// (1), same final stdout line. Parse, solve and print go through the C ABI of
// libkc_amd.so. KC_VERBOSE=1 adds stage timings on stderr.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/ktruss.h"

int main(int argc, char* argv[]) {
  if (argc != 3 || std::strcmp(argv[1], "-f") != 0) {
    std::fprintf(stderr, "Usage: -f [data_file_path]\n");
    return 1;
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
  return 0;
}
