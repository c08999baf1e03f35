// Launch wrappers for the device stages of the k-max-truss path.
// Every wrapper enqueues on `s` and returns immediately.
#pragma once
#include "kt_common.h"

namespace kt {

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);

// (a) undirected CSR from m canonical pairs (both directions inserted) or
//     from L directed lines (x into row y). deg must be zeroed first.
void csr_degree_pairs(const uint32_t* pairs, uint64_t m, uint32_t n, uint32_t* deg, Ctl* ctl,
                      hipStream_t s);
void csr_degree_lines(const uint32_t* xs, const uint32_t* ys, uint64_t L, uint32_t n,
                      uint32_t* deg, Ctl* ctl, hipStream_t s);
void csr_scatter_pairs(const uint32_t* pairs, uint64_t m, uint32_t n, uint64_t* cursor,
                       uint32_t* adj, hipStream_t s);
void csr_scatter_lines(const uint32_t* xs, const uint32_t* ys, uint64_t L, uint32_t n,
                       uint64_t* cursor, uint32_t* adj, hipStream_t s);
void degrees_from_offsets(const uint64_t* off, uint32_t n, uint32_t* deg, hipStream_t s);

// (b) vertex peeling (k-core). lvl[v] == kNone: alive; otherwise the
//     threshold at which v was removed. deg holds alive-neighbour counts.
void core_frontier_init(const uint32_t* lvl, const uint32_t* deg, uint32_t n, uint32_t k,
                        uint32_t* q, uint32_t* qcount, hipStream_t s);
void core_stamp(const uint32_t* q, const uint32_t* qcount, uint32_t maxq, uint32_t k,
                uint32_t* lvl, hipStream_t s);
void core_process(const uint64_t* off, const uint32_t* adj, const uint32_t* q,
                  const uint32_t* qcount, uint32_t maxq, uint32_t k, uint32_t* lvl,
                  uint32_t* deg, uint32_t* qnext, uint32_t* qnext_count, hipStream_t s);
void core_alive_degrees(const uint64_t* off, const uint32_t* adj, const uint32_t* lvl,
                        uint32_t n, uint32_t* deg, hipStream_t s);
// lvl[v] = (lvl[v] > k) ? kNone : (lvl[v] == kNone ? kNone : k) -- used to
// re-open every vertex whose removal threshold exceeds k.
void core_reopen_above(uint32_t* lvl, uint32_t n, uint32_t k, hipStream_t s);

// (a1-a3) core subgraph extraction.
void core_collect(const uint32_t* lvl, const uint32_t* deg, uint32_t n, uint64_t* keys,
                  uint32_t* count, hipStream_t s);
void core_assign_rank(const uint64_t* sorted_keys, uint32_t nc, uint32_t* to, uint32_t* inv,
                      hipStream_t s);
void orient_count(const uint64_t* off, const uint32_t* adj, const uint32_t* to,
                  const uint32_t* inv, uint32_t nc, uint32_t* dout, hipStream_t s);
void orient_fill(const uint64_t* off, const uint32_t* adj, const uint32_t* to,
                 const uint32_t* inv, uint32_t nc, const uint32_t* off_d, uint32_t* adj_d,
                 uint32_t* src, uint32_t* din, hipStream_t s);
// Sort every row of a CSR in place (ascending). Rows longer than
// `lds_cap` elements are left for the caller's fallback; returns nothing.
void sort_rows(const uint32_t* off, uint32_t nrows, uint32_t* vals, uint32_t max_row,
               hipStream_t s);
void build_undirected(const uint32_t* off_d, const uint32_t* adj_d, const uint32_t* din_off,
                      const uint32_t* in_src_sorted, const uint32_t* in_eid_sorted, uint32_t nc,
                      const uint32_t* off_u, uint32_t* adj_u, uint32_t* eid_u, hipStream_t s);
void iota_u32(uint32_t* p, uint32_t n, hipStream_t s);
void add_offsets(const uint32_t* a, const uint32_t* b, uint32_t n, uint32_t* out, hipStream_t s);
void gather_u32(const uint32_t* idx, const uint32_t* table, uint32_t n, uint32_t* out,
                hipStream_t s);

// (c) supports, plain (not doubled) int32 per oriented edge; sup zeroed first.
// `lanes` selects the sub-wavefront width per row.
bool support_bitmap_fits(uint32_t nc, uint32_t max_dout);
void support_bitmap(const uint32_t* off_d, const uint32_t* adj_d, uint32_t nc, uint32_t max_dout,
                    int lanes, int32_t* sup, Ctl* ctl, hipStream_t s);
void support_edges(const uint32_t* off_d, const uint32_t* adj_d, const uint32_t* src, uint32_t nc,
                   uint32_t mc, int lanes, int32_t* sup, hipStream_t s);

// (d) edge peeling. state[e] == kNone: alive; r: in round r's frontier;
//     < r: deleted earlier.
void edge_frontier_init(const uint32_t* state, const int32_t* sup, uint32_t mc, int32_t thr,
                        uint32_t* q, uint32_t* qcount, hipStream_t s);
void edge_stamp(const uint32_t* q, const uint32_t* qcount, uint32_t maxq, uint32_t r,
                uint32_t* state, hipStream_t s);
void peel_process(const uint32_t* off_u, const uint32_t* adj_u, const uint32_t* eid_u,
                  const uint32_t* src, const uint32_t* dst, const uint32_t* q,
                  const uint32_t* qcount, uint32_t maxq, uint32_t r, int32_t thr, int lanes,
                  uint32_t* state, int32_t* sup, uint32_t* qnext, uint32_t* qnext_count,
                  hipStream_t s);
void edge_min_support(const uint32_t* state, const int32_t* sup, uint32_t mc, uint32_t* out,
                      hipStream_t s);
void edges_alive_above(const uint32_t* state, uint32_t mc, uint32_t stamp_floor,
                       const uint32_t* src, const uint32_t* dst, const uint32_t* inv,
                       uint32_t* pairs_out, uint32_t* count, hipStream_t s);

void w_stats(const uint32_t* off_d, const uint32_t* src, const uint32_t* dst, const uint32_t* state,
             uint32_t mc, uint32_t floor, uint64_t* out2, hipStream_t s);

}  // namespace kt
