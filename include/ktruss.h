/*
 * ktruss.h — C ABI of the MI355X k-max-truss library (libkc_amd.so).
 *
 * Plain pointers and sizes only. Every function returns 0 on success and a
 * non-zero status on failure; kt_last_error() then describes the failure.
 * A context owns one HIP stream and its device buffers; calls on one context
 * must not overlap (the same holds for the reference, whose module-level
 * device counters make it non-reentrant, src/kc.cu:55-58).
 *
 * Which reference interface each entry point replaces:
 *   kt_solve_tsv            the whole `kc_gpu -f <file>` run: parse + two-pass
 *                           solve + result line (src/kc_gpu.cpp:209-347)
 *   kt_solve_lines_host     kmax_truss() called twice on the parsed graph
 *                           (src/kc_gpu.cpp:129-208, 336-341)
 *   kt_solve_pairs_host     same, from canonical undirected pairs
 *   kt_solve_pairs_device   same, input already resident in HBM (no reference
 *                           counterpart: the reference copies host arrays in
 *                           every call, src/kc.cu:584-651)
 *   kt_truss_edges          the kmax-truss edge set, which the reference
 *                           computes but never emits (src/kc.cu:1236-1242)
 * The per-stage functions of src/kc_gpu.h are declared in kc_gpu.h (C++ ABI,
 * same signatures) and in kc_gpu_c.h (C ABI).
 */
#ifndef KT_KTRUSS_H
#define KT_KTRUSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kt_context kt_context;

typedef struct kt_result {
  uint32_t kmax;          /* largest k with a non-empty k-truss; 0 if triangle-free */
  uint32_t max_core;      /* largest k with a non-empty k-core */
  uint64_t truss_edges;   /* edges in the kmax-truss */
  uint32_t n;             /* vertex count of the input */
  uint32_t k_hat;         /* pass 1: kmax of the max-core (a lower bound) */
  uint64_t m;             /* undirected edges of the input */
  uint64_t k_hat_edges;
  uint32_t core_k;        /* pass 2 works on the core_k-core */
  uint32_t n_core;
  uint64_t m_core;
  uint64_t peel_rounds;   /* edge-peel rounds, both passes */
  uint32_t levels;        /* non-empty truss levels visited */
  uint32_t lanes;         /* sub-wavefront width chosen for the support count */
  float ms_total;         /* device time of the whole solve */
  float ms_build;         /* (a) CSR build */
  float ms_kcore;         /* (b) vertex peeling, both passes */
  float ms_pass1;         /* pass 1 on the max-core */
  float ms_extract;       /* (a1-a3) pass-2 core extraction */
  float ms_count;         /* (c) pass-2 support count */
  float ms_peel;          /* (d) pass-2 edge peeling */
  float ms_count_kernel;  /* device time of the pass-2 support kernel alone */
  uint32_t compactions;   /* edge-set compactions, both passes (src/kc.cu:1059-1138) */
  uint32_t recounts;      /* dropped frontiers recounted exactly (src/kc.cu:1206-1233) */
} kt_result;

/* Byte-model terms of the pass-2 core of the last solve: C is the
 * core_k-core, C+ its rank orientation, W(S) = sum over oriented edges of S
 * of d+(u) + d+(v), T the kmax-truss. */
typedef struct kt_stats {
  uint32_t n_core;
  uint32_t max_out_degree;
  uint64_t m_core;
  uint64_t w_core;      /* W(C+) */
  uint64_t w_peeled;    /* W(C+ \ T+) */
  uint64_t truss_edges;
  uint32_t count_kernel;  /* pass-2 support count: 0 LDS bitmap, 1 per edge, 2 bit matrix */
  uint32_t core_words;    /* bit-matrix row words (count_kernel 2) */
  uint32_t vertex_rounds; /* the last vertex peel: 0 one cascade launch, 1 synchronous rounds */
  uint32_t pad;
} kt_stats;

/* Library and device */
const char* kt_last_error(void);
int kt_device_count(int* count);
int kt_create(int device, kt_context** out);
int kt_destroy(kt_context* ctx);

/* Device memory helpers for callers without their own allocator. */
int kt_malloc(kt_context* ctx, uint64_t bytes, void** dptr);
int kt_free(kt_context* ctx, void* dptr);
int kt_memcpy_h2d(kt_context* ctx, void* dst, const void* src, uint64_t bytes);
int kt_memcpy_d2h(kt_context* ctx, void* dst, const void* src, uint64_t bytes);
int kt_memcpy_d2d(kt_context* ctx, void* dst, const void* src, uint64_t bytes);
int kt_synchronize(kt_context* ctx);

/* Whole solves. Pairs are interleaved (u0, v0, u1, v1, ...), u != v, each
 * undirected edge once, ids < n. Lines follow the reference's input
 * contract: line i puts xs[i] into row ys[i]; both directions present. */
int kt_solve_pairs_device(kt_context* ctx, const uint32_t* d_pairs, uint64_t m, uint32_t n,
                          kt_result* res);
int kt_solve_pairs_host(kt_context* ctx, const uint32_t* pairs, uint64_t m, uint32_t n,
                        kt_result* res);
int kt_solve_lines_host(kt_context* ctx, const uint32_t* xs, const uint32_t* ys, uint64_t L,
                        uint32_t n, kt_result* res);
/* Edge-list text file (src/kc_gpu.cpp:243-312 parses it on the host): the
 * bytes are streamed into HBM and parsed there (a file with a line holding a
 * single number goes through the host parser instead, whose reading of such
 * a line spans the next one). parse_ms: file to line arrays in HBM. */
int kt_solve_tsv(kt_context* ctx, const char* path, kt_result* res, double* parse_ms);
/* The device-side parse alone, lines copied back (at most cap of each);
 * *on_device is 0 when the host parser had to take the file. */
int kt_parse_tsv_device(kt_context* ctx, const char* path, uint32_t* xs, uint32_t* ys, uint64_t cap,
                        uint64_t* lines, uint32_t* n, int* on_device);

/* Survivors of the last solve: original ids, u < v, unordered.
 * *count receives the truss size; at most cap pairs are written. */
int kt_truss_edges(kt_context* ctx, uint32_t* pairs, uint64_t cap, uint64_t* count);

/* Per-edge trussness of the last solved graph (no counterpart in the
 * reference, which reports kmax and |T| only): a
 * full truss decomposition on the device, then (u, v, t) triples, u < v in
 * original ids, for every edge of trussness t >= 3 (every other edge has
 * trussness 2), unordered. *count receives how many; at most cap triples
 * are written. The decomposition runs once per input graph: a second call
 * (e.g. with a buffer sized by the first call's count) only copies. The last
 * solve's kmax-truss stays available to kt_truss_edges; kt_last_stats fails
 * afterwards (the decomposition reuses the core buffers it describes). */
int kt_trussness(kt_context* ctx, uint32_t* triples, uint64_t cap, uint64_t* count);

/* Host TSV parser (the reference's input contract, src/kc_gpu.cpp:243-312).
 * Pass NULL buffers to get the line count and n only. */
int kt_parse_tsv(const char* path, uint32_t* xs, uint32_t* ys, uint64_t cap, uint64_t* lines,
                 uint32_t* n);

/* Tuning: sub-wavefront lanes per row for the support count and the peel
 * (0 = automatic; 8, 16, 32 or 64) and kernel-variant flags:
 * KT_FORCE_EDGE_COUNT selects the per-edge support kernel instead of the LDS
 * bitmap one, KT_FORCE_ATOMIC_BUILD the atomic CSR build instead of the
 * bucketed one, KT_FORCE_UNSORTED_BUILD the bucketed build even for inputs
 * sorted by row (default: lines sorted by row are read as runs, and pairs
 * sorted by their first id have only their second entries bucketed),
 * KT_FORCE_SPARSE / KT_FORCE_DENSE the sorted-row or the
 * bit-matrix core representation (default: by density), KT_FORCE_LEVEL_WALK
 * the level-by-level first pass on bit-matrix cores (default: 16 levels are
 * walked, then thresholds are probed), KT_FORCE_PROBES threshold probes right
 * after the first level, KT_FORCE_VERTEX_ROUNDS / KT_FORCE_VERTEX_CASCADE
 * the vertex peel by synchronous rounds / by one cascade launch per
 * threshold at every level (default: the cascade on levels of up to 2^24
 * row slots, rounds above, and the max-core search finished in one
 * workgroup on levels of up to 16,000 vertices, which either flag turns
 * off). Every variant computes the same result. The
 * sparse-peel drop ratio returns to its default (KT_SPARSE_DROP_RATIO or
 * 3). */
#define KT_FORCE_EDGE_COUNT 1
#define KT_FORCE_ATOMIC_BUILD 2
#define KT_FORCE_SPARSE 4
#define KT_FORCE_DENSE 8
#define KT_FORCE_LEVEL_WALK 16
#define KT_FORCE_PROBES 32
#define KT_FORCE_VERTEX_ROUNDS 64
#define KT_FORCE_VERTEX_CASCADE 128
#define KT_FORCE_UNSORTED_BUILD 256
int kt_set_tuning(kt_context* ctx, int count_lanes, int peel_lanes, int flags);

/* Batched solve of `count` independent graphs (no counterpart in the
 * reference, which solves one graph per process, src/kc_gpu.cpp:209-347):
 * their canonical pairs (u < v < n, strictly sorted within each graph) lie
 * concatenated in HBM at d_pairs; h_offsets (count + 1 pair offsets) and h_ns
 * (vertex counts) are host arrays. Per graph: d_kmax, d_count (edges in its
 * kmax-truss) and d_status (0 ok, 1 invalid input); d_keep (nullable) gets 1
 * for every pair in its graph's kmax-truss. The valid graphs are solved
 * together as disjoint unions: the batch is cut into up to four contiguous
 * parts of about equal edges (KT_BATCH_PARTS sets the count), and each
 * part's union is decomposed on its own solver and stream (the context's
 * for the first), all at once; every graph's answer is read off its part's
 * decomposition. Results equal kt_solve_pairs_* on each graph. flags is
 * reserved and must be 0. */
int kt_batch_solve(kt_context* ctx, const uint32_t* d_pairs, const uint64_t* h_offsets,
                   const uint32_t* h_ns, uint32_t count, uint32_t* d_kmax, uint64_t* d_count,
                   uint32_t* d_status, uint8_t* d_keep, int flags);
/* Autotuner of the batched solve (no counterpart in the reference): times
 * kt_batch_solve on every sample_every-th graph of the batch (1: all) under
 * candidate settings, one at a time: sparse-peel lanes, support-count lanes,
 * support-count kernel, drop ratio (a candidate is kept when it is 3%
 * faster than the best so far). Every trial must give every graph the same
 * answer (else an error is returned). The fastest settings stay on the
 * context for later solves (kt_set_tuning resets them) and the whole batch is
 * then solved into d_kmax / d_count / d_status as kt_batch_solve does. */
typedef struct kt_batch_tuning {
  int32_t count_lanes;  /* support count: sub-wavefront lanes per row (0 automatic) */
  int32_t peel_lanes;   /* sparse peel rounds: the same */
  int32_t edge_count;   /* 1: per-edge support kernel (as KT_FORCE_EDGE_COUNT) */
  uint32_t trials;      /* batch solves timed */
  double drop_ratio;    /* sparse peel: recount when frontier * ratio >= live edges */
  double ms_default;    /* sampled batch time with the context's settings on entry */
  double ms_best;       /* with the chosen settings */
} kt_batch_tuning;
int kt_batch_autotune(kt_context* ctx, const uint32_t* d_pairs, const uint64_t* h_offsets,
                      const uint32_t* h_ns, uint32_t count, uint32_t sample_every, uint32_t* d_kmax,
                      uint64_t* d_count, uint32_t* d_status, kt_batch_tuning* out);

/* Serving (no counterpart in the reference, which answers one graph per
 * process, src/kc_gpu.cpp:209-347): queries are queued first come, first
 * served and solved concurrently on `slots` solver contexts, each with its
 * own HIP stream and host thread. Each slot keeps its CSR, support and peel
 * buffers resident between queries and is charged against hbm_budget what
 * they hold, plus the footprint of its running query (kt_serve_footprint).
 * A query is admitted as soon as a slot is free and that charge fits in the
 * budget; when nothing runs and it still does not fit, idle slots free their
 * buffers. A query that could not fit even then is rejected (status 1).
 * Continuous batching: when more small queries wait than idle slots can
 * take, a slot takes several (each at most 2^16 edges, up to 64 queries and
 * 2^22 edges in all, each one charged its footprint) and solves them as one
 * disjoint union, as kt_batch_solve does. kmax and truss_edges equal
 * kt_solve_pairs_* on the same graph, alone or in the mix; a grouped query's
 * result carries n, m, kmax and truss_edges, its other fields are 0. */
typedef struct kt_server kt_server;
typedef struct kt_serve_result {
  int32_t status;         /* 0 solved, 1 rejected by admission, 2 solve failed */
  uint32_t slot;          /* slot that solved it */
  kt_result result;
  uint64_t footprint;     /* kt_serve_footprint of the query */
  uint64_t charged;       /* bytes its slot was charged while it ran */
  uint64_t slot_bytes;    /* device bytes its slot held after the solve */
  double queued_ms;       /* submit -> admitted */
  double run_ms;          /* admitted -> solved, upload included */
  double total_ms;        /* submit -> solved */
  uint32_t group;         /* queries solved with it as one disjoint union (1: alone) */
} kt_serve_result;
typedef struct kt_serve_counters {
  uint64_t submitted, completed, rejected, failed;
  uint64_t in_flight, peak_in_flight;
  uint64_t bytes_in_use, peak_bytes_in_use;
  uint64_t over_estimate; /* solves whose slot grew past the bytes charged for them */
  uint64_t trims;         /* idle slots that freed their buffers to let a query in */
  uint64_t grouped;       /* queries solved in a union with others */
} kt_serve_counters;
int kt_serve_create(int device, int slots, uint64_t hbm_budget, kt_server** out);
int kt_serve_destroy(kt_server* srv);
/* pairs: canonical pairs (u < v < n) in host memory, or device memory when
 * on_device is 1; they must stay valid until kt_serve_wait returns for the
 * ticket. */
int kt_serve_submit(kt_server* srv, const uint32_t* pairs, uint64_t m, uint32_t n, int on_device,
                    uint64_t* ticket);
/* Blocks until the query is done; its error text (status 2) is then
 * kt_last_error(). Each ticket can be waited for once. */
int kt_serve_wait(kt_server* srv, uint64_t ticket, kt_serve_result* out);
int kt_serve_counters_get(kt_server* srv, kt_serve_counters* out);
/* Admission charge of a query of m edges on n vertices (bytes). */
uint64_t kt_serve_footprint(uint64_t m, uint32_t n);

/* Measurement hook for the vertex-peel push kernel (k_core_push), the
 * largest kernel of a solve. mode bit 1 (KT_PROBE_TIME): each launch stamps
 * HIP events at its own start and end on the context's stream; bit 2
 * (KT_PROBE_COUNT): each launch counts the row slots it walks (rows up to
 * 8192 slots; longer ones go to the sliced kernel) and the vertices it takes,
 * with atomics that slow it, so time and count in separate solves. 0 turns
 * the probe off. kt_push_stats_get returns the sums since the last
 * kt_push_probe. */
#define KT_PROBE_TIME 1
#define KT_PROBE_COUNT 2
typedef struct kt_push_stats {
  uint64_t launches;
  uint64_t slots;
  uint64_t vertices;
  double ms;  /* summed launch durations */
} kt_push_stats;
int kt_push_probe(kt_context* ctx, int mode);
int kt_push_stats_get(kt_context* ctx, kt_push_stats* out);
/* The same measurement for another kernel: KT_PROBE_VERTEX_PEEL is what
 * kt_push_probe measures (the vertex-peel cascade launch, or the push of
 * the round-based peel); KT_PROBE_DENSE_PEEL is the bit-matrix edge-peel
 * round (k_dense_peel_round), whose counters in kt_push_stats are slots =
 * the frontier edges' oriented out-degrees d+(u) + d+(v) summed, vertices =
 * the frontier edges walked. */
#define KT_PROBE_VERTEX_PEEL 0
#define KT_PROBE_DENSE_PEEL 1
int kt_kernel_probe(kt_context* ctx, int kernel, int mode);

/* Byte-model terms of the last solve (one extra device pass, untimed). */
int kt_last_stats(kt_context* ctx, kt_stats* out);

/* Stage access for parity tests: the undirected CSR built by the last solve
 * (off has n+1 entries, adj 2m). */
int kt_last_csr(kt_context* ctx, uint64_t* off, uint32_t* adj);

/* Allocation-free solve. No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The same solve and
 * results as kt_solve_pairs_device, but every device buffer it needs is
 * carved from a fixed scratch array of KT_NOALLOC_SCRATCH_BYTES per GPU (a
 * file-scope __device__ array of the library, held on each GPU it runs on
 * until the process ends); no context is needed and no hipMalloc, hipFree
 * or allocating library routine is called. The first call on a device sets
 * up that device's solver (its stream and timing events). A graph is
 * admitted when 160*m + 64*n + min(72*m, 32 MiB) + min(512*m, 32 MiB) +
 * 16 MiB <= KT_NOALLOC_SCRATCH_BYTES - 4 KiB (about 1.9M edges on few
 * vertices); larger ones are rejected with an error before any work. When
 * d_truss is non-null, the kmax-truss pairs (input ids, u < v, unordered;
 * at most cap of them) are written there (device memory, 2*cap uint32);
 * *truss_count (nullable) receives the truss size. Calls on one device are
 * serialised in the process; calls from several processes on one device
 * share its array and race. Returns when the results are written. */
#define KT_NOALLOC_SCRATCH_BYTES (512ull << 20)
int kt_solve_pairs_device_noalloc(int device, const uint32_t* d_pairs, uint64_t m, uint32_t n,
                                  uint32_t* d_truss, uint64_t cap, uint64_t* truss_count,
                                  kt_result* res);

#ifdef __cplusplus
}
#endif
#endif
