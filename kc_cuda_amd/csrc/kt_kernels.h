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

// Partitioned build: entries bucketed by row >> S (B buckets, T tiles of
// `tile` items), then rows built per bucket in LDS. mode 0: a = interleaved
// pairs (both directions); mode 1: a = xs (columns), b = ys (rows); mode 2:
// pairs, second entries (row v, column u) only.
// hist_t has T*B entries (tile-major, one row per tile); its transpose is
// scanned bucket-major into hoff (B*T+1), whose tile-major transpose hoff_t
// gives each tile its write cursors.
bool csr_partition_plan(uint32_t n, uint32_t* S, uint32_t* B);
void csr_part_hist(int mode, const uint32_t* a, const uint32_t* b, uint64_t items, uint64_t tile,
                   uint32_t n, uint32_t S, uint32_t B, uint32_t T, uint32_t* hist_t, uint32_t* err,
                   hipStream_t s);
void transpose_u32(const uint32_t* in, uint32_t rows, uint32_t cols, uint32_t* out, hipStream_t s);
void transpose_u64(const uint64_t* in, uint32_t rows, uint32_t cols, uint64_t* out, hipStream_t s);
void csr_part_scatter(int mode, const uint32_t* a, const uint32_t* b, uint64_t items, uint64_t tile,
                      uint32_t n, uint32_t S, uint32_t B, uint32_t T, const uint64_t* hoff_t,
                      uint32_t* part, hipStream_t s);
// fwd_pairs / f0 (row-sorted pairs, else null): the pairs' first entries,
// read in place (f0 from csr_fwd_start); part then holds only their second
// entries (mode 2 of csr_part_hist / csr_part_scatter).
void csr_bucket_rows(const uint32_t* part, const uint64_t* hoff, const uint32_t* fwd_pairs, const uint64_t* f0,
                     uint32_t B, uint32_t T, uint32_t S, uint32_t n, uint64_t* off, uint32_t* deg, uint32_t* adj,
                     hipStream_t s);
// Row-sorted inputs. csr_sorted_check: *unsorted |= 1 unless the rows are
// non-decreasing (mode 0: pairs by first id; mode 1: lines by ys); for lines
// also hpos[r] / tpos[r] = first / one past the last line of row r (fill
// both with 0 first) and *err |= 1 for a row >= n. csr_deg_from_runs:
// deg[0..n] from them (deg[n] = 0). csr_copy_cols: adj = xs (*err |= 1 for a
// column >= n). csr_fwd_start: f0[b] = first pair with first id >= b << S,
// f0[B] = items.
void csr_sorted_check(int mode, const uint32_t* keys, uint64_t items, uint32_t n, uint64_t* hpos, uint64_t* tpos,
                      uint32_t* unsorted, uint32_t* err, hipStream_t s);
void csr_deg_from_runs(const uint64_t* hpos, const uint64_t* tpos, uint32_t n, uint32_t* deg, hipStream_t s);
void csr_copy_cols(const uint32_t* xs, uint64_t L, uint32_t n, uint32_t* adj, uint32_t* err, hipStream_t s);
void csr_fwd_start(const uint32_t* pairs, uint64_t items, uint32_t S, uint32_t B, uint64_t* f0, hipStream_t s);

// (b) vertex peeling (k-core). lvl[v] == kNone: alive; otherwise the
//     threshold at which v was removed. deg holds alive-neighbour counts.
// Device-driven rounds: each core_round queues and stamps (lvl = k) every
// alive vertex of degree < k, then takes one degree off every neighbour of
// the queued vertices. ctl->removed counts the removed vertices;
// ctl->action turns idle after a round that found nothing to remove.
// q holds n entries. Measurement (all nullable): push_stats[0] / [1] gain
// the light row slots the push kernel walks and the vertices it takes; ev0
// and ev1 (both or neither) are stamped at its start and end.
void core_round_start(Ctl* ctl, hipStream_t s);
void core_round(const uint64_t* off, const uint32_t* adj, uint32_t n, uint32_t* q, Ctl* ctl, uint32_t k,
                uint32_t* lvl, uint32_t* deg, HeavyList hl, unsigned long long* push_stats, hipEvent_t ev0,
                hipEvent_t ev1, hipStream_t s);
// The same peel at one threshold k as a single cascade launch (a work queue
// of vertices, each queued once: by the initial sweep of degree < k or when a
// decrement takes its degree from k to k - 1). q holds cap 8-byte items
// (cascade_capacity of the level); afterwards ctl->vq_reserve - vq_chunks
// vertices were removed (lvl = k), and ctl->vq_err is non-zero if the queue
// overflowed or a wave gave up waiting. Measurement as for core_round.
uint64_t cascade_capacity(uint32_t n, uint64_t slots);
// stop_at != 0: leave as soon as stop_at vertices are queued (the k-core is
// then empty; ctl->vq_err gets kVqEmpty and the level is left part-peeled).
void core_cascade(const uint64_t* off, const uint32_t* adj, uint32_t n, uint32_t k, uint32_t* lvl, uint32_t* deg,
                  unsigned long long* q, uint32_t cap, Ctl* ctl, unsigned long long* stats, hipEvent_t ev0,
                  hipEvent_t ev1, hipStream_t s, uint32_t stop_at = 0);
void core_alive_degrees(const uint64_t* off, const uint32_t* adj, const uint32_t* lvl,
                        uint32_t n, uint32_t* deg, HeavyList hl, hipStream_t s);
// lvl[v] = (lvl[v] > k) ? kNone : (lvl[v] == kNone ? kNone : k) -- used to
// re-open every vertex whose removal threshold exceeds k.
void core_reopen_above(uint32_t* lvl, uint32_t n, uint32_t k, hipStream_t s);
// Max core of a level of at most core_decompose_max_n() vertices in one
// workgroup, from its current state (the alive vertices, their alive-
// neighbour degrees, every alive degree >= k0 - 1): thresholds walked upward
// from k0 without restores; afterwards the max core is alive with its
// degrees, every other vertex v of the level removed here has lvl[v] = its
// core number + 1, out[0] = the max core's k (0 if nothing was alive) and
// out[1] its vertex count. q holds n entries.
uint32_t core_decompose_max_n();
void core_decompose(const uint64_t* off, const uint32_t* adj, uint32_t n, uint32_t k0, uint32_t* lvl, uint32_t* deg,
                    uint32_t* q, uint32_t* out, hipStream_t s);
// Degree histogram (kDegBins bins: exact below kDegExact, then one per power
// of two) and the degree filter of the max-core search (lvl = k for every
// vertex of degree < k, kNone otherwise; *alive += vertices kept).
constexpr uint32_t kDegExactLog2 = 12;
constexpr uint32_t kDegExact = 1u << kDegExactLog2;
constexpr uint32_t kDegBins = kDegExact + 32 - kDegExactLog2;
void degree_hist(const uint32_t* deg, uint32_t n, uint32_t* hist, hipStream_t s);
void core_filter(const uint32_t* deg, uint32_t n, uint32_t k, uint32_t* lvl, uint32_t* alive, hipStream_t s);

// (a1-a3) core subgraph extraction.
// key = degree << 32 | ~original id (to_orig may be null: identity), val = local id
void core_collect(const uint32_t* lvl, const uint32_t* deg, const uint32_t* to_orig, uint32_t n,
                  uint64_t* keys, uint32_t* vals, uint32_t* count, hipStream_t s);
void core_assign_rank(const uint64_t* sorted_keys, const uint32_t* sorted_vals, uint32_t nc,
                      uint32_t* to, uint32_t* inv, uint32_t* inv_orig, hipStream_t s);
// (b') induced subgraph of the alive vertices
void alive_list(const uint32_t* lvl, uint32_t n, uint32_t* list, uint32_t* new_id,
                uint32_t* count, hipStream_t s);
// sub_deg[i] = deg[list[i]] (skipped when deg is null), sub_orig[i] = its input id.
void sub_degrees(const uint32_t* list, const uint32_t* deg, const uint32_t* to_orig, uint32_t na,
                 uint32_t* sub_deg, uint32_t* sub_orig, hipStream_t s);
// sub_deg[i] = neighbours of list[i] whose bit is set in the alive bitmap
// abits (bit w % 32 of word w / 32).
void list_alive_degrees(const uint64_t* off, const uint32_t* adj, const uint32_t* list, uint32_t na,
                        const uint32_t* abits, const uint32_t* acoarse, uint32_t* sub_deg, HeavyList hl,
                        hipStream_t s);
// coarse: one bit per bitmap word ((n + 31) / 32 words -> ((n + 31) / 32 + 31) / 32 entries), set when
// the word has an alive vertex; the cut's lookups read bits and pref only under a set bit.
void alive_coarse(const uint32_t* bits, uint32_t n, uint32_t* coarse, hipStream_t s);
// Alive vertices of a level as a bitmap (bits: (n + 31) / 32 words) with the
// popcount of every word in cnt (words + 1 entries, the last set to 0); an
// exclusive scan of cnt gives pref, and ranked_list writes the alive vertices
// in increasing id (list[rank] = vertex).
void alive_words(const uint32_t* lvl, uint32_t n, uint32_t* bits, uint32_t* cnt, hipStream_t s);
void alive_ranked_list(const uint32_t* bits, const uint32_t* pref, uint32_t n, uint32_t* list, hipStream_t s);
// Rows of the cut: each alive neighbour w becomes its rank (abits/apref from
// alive_words and the scan). cursor: cursor_cap zeroed 64-bit slots, at least
// the number of rows of the cut longer than the heavy threshold (slots / 8192
// + 1 suffices)
void sub_fill(const uint64_t* off, const uint32_t* adj, const uint32_t* list, const uint32_t* abits,
              const uint32_t* apref, const uint32_t* acoarse, uint32_t na, const uint64_t* sub_off, uint32_t* sub_adj,
              HeavyList hl, uint64_t* cursor, uint32_t cursor_cap, hipStream_t s, uint32_t* sub_deg = nullptr);
// The cut without a counting pass: len[0..na] = the listed rows' lengths
// (len[na] = 0), scanned into upper-bound offsets for sub_fill (with
// sub_deg: it then counts each row's alive neighbours), then rows_pack
// moves the rows to their exact offsets.
void list_row_lengths(const uint64_t* off, const uint32_t* list, uint32_t na, uint32_t* len, hipStream_t s);
void rows_pack(const uint32_t* tmp, const uint64_t* ub_off, const uint64_t* off, const uint32_t* deg, uint32_t na,
               uint32_t* adj, hipStream_t s);
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
void edge_min_support(const uint32_t* state, const int32_t* sup, uint32_t mc, uint32_t* out,
                      hipStream_t s);

void w_stats(const uint32_t* off_d, const uint32_t* src, const uint32_t* dst, const uint32_t* del_thr,
             uint32_t mc, int32_t floor, uint64_t* out2, hipStream_t s);

// Dense cores: neighbourhoods as bit-matrix rows of ws = ceil(nc / 64) words.
uint32_t dense_words(uint32_t nc);
bool dense_fits(uint32_t nc, uint32_t mc);
void dense_rows(const uint32_t* off_u, const uint32_t* adj_u, uint32_t nc, uint32_t ws,
                uint64_t* bits0, uint64_t* bits, uint32_t* pref0, hipStream_t s);
void dense_support(const uint32_t* off_d, const uint32_t* dst, uint32_t nc, uint32_t mc, const uint64_t* bits,
                   uint32_t ws, const uint32_t* state, int32_t* sup, hipStream_t s);

// Device-driven edge-peel rounds and truss levels (kt_peel.hip). q0/q1 are
// the two frontier queues; ctl->cur says which one holds the current
// frontier, ctl->thr the level threshold, ctl->action what comes next.
// Starts a frontier already collected into q0 (count in ctl->qcount[0]) at
// threshold thr with `alive` live edges; reset also starts a new search
// (level record cleared, search rule set). Stamps nothing.
void round_start(Ctl* ctl, uint32_t r, uint64_t alive, int32_t thr, const RoundRule& rule, bool reset,
                 const SearchRule& search, hipStream_t s);
void round_stamp(const PeelArgs& a, Ctl* ctl, const RoundRule& rule, hipStream_t s);
// `count` rounds of (peel, advance, stamp)
void peel_rounds(const PeelArgs& a, Ctl* ctl, const RoundRule& rule, int lanes, int count, hipStream_t s,
                 const LaunchProbe& probe = LaunchProbe{});
void frontier_clear(const PeelArgs& a, const Ctl* ctl, hipStream_t s);
void dense_eidm(const uint32_t* off_u, const uint32_t* adj_u, const uint32_t* eid_u, uint32_t nc,
                uint32_t* eidm, hipStream_t s);
// Vertex compaction of a core view: flag the endpoints of live edges (flag
// zeroed first), then, with vmap = exclusive scan of the flags, renumber the
// compacted edges and rebuild the row offsets (off gets vmap[n] + 1 entries).
void live_vertex_flags(const uint32_t* state, const uint32_t* src, const uint32_t* dst, uint32_t mc,
                       uint32_t* flag, hipStream_t s);
void relabel_edges(uint32_t* adj, uint32_t* src, uint32_t m, const uint32_t* vmap, hipStream_t s);
void relabel_offsets(const uint32_t* roff, const uint32_t* vmap, uint32_t n, uint32_t* off, hipStream_t s);

void row_src(const uint32_t* off, uint32_t rows, uint32_t* src, hipStream_t s);
void sons(const uint32_t* off_u, const uint32_t* adj_u, const uint32_t* eid_u, uint32_t nc,
          uint32_t* lson, uint32_t* rson, hipStream_t s);

// (d3) compaction of the live edges and deletion bookkeeping on base ids
void row_alive_count(const uint32_t* off_d, const uint32_t* state, uint32_t nc, uint32_t* cnt,
                     hipStream_t s);
void row_compact(const uint32_t* off_d, const uint32_t* adj_d, const uint32_t* orig,
                 const int32_t* sup, const uint32_t* state, uint32_t nc, const uint32_t* off2,
                 uint32_t* adj2, uint32_t* src2, uint32_t* orig2, int32_t* sup2, hipStream_t s);
void count_in_degrees(const uint32_t* adj, uint32_t m, uint32_t* din, hipStream_t s);
void truss_pairs(const uint32_t* del_thr, uint32_t m0, int32_t floor, const uint32_t* src0,
                 const uint32_t* dst0, const uint32_t* inv, uint32_t* out, uint32_t* count,
                 hipStream_t s);

}  // namespace kt
