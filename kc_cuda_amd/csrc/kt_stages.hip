// Stage logic of the device-resident two-pass k-max-truss solve (host side).
//
// Flow (reference: src/kc_gpu.cpp:129-208, 336-341; src/kc.cu:770-884, 1140-1309):
//   (b) vertex peel with a galloping threshold -> max-core C0. When a step
//       leaves few vertices alive, the rest of the peel continues on their
//       induced subgraph (a new level with dense ids), so high thresholds
//       never rescan the rows of vertices already gone.
//   pass 1: rank + orient C0, exact supports, truss levels -> k̂ <= kmax.
//   pass 2: the (L-1)-core C, L = max(k̂, 3) (it holds the kmax-truss),
//       recovered from the deepest level that contains it; rank + orient,
//       exact supports, truss levels upward from L with the reference's early
//       exit (first level equal to pass 1's truss, src/kc.cu:1240-1241).
//   Peeling drops a frontier holding a third of the live edges and recounts
//   the rest (src/kc.cu:988, 1206-1233), and compacts once half the edges
//   are dead, carrying supports (src/kc.cu:1174-1178).
// Only scalar counters cross to the host.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "kt_batch_union.h"
#include "kt_kernels.h"
#include "kt_solver.h"

namespace kt {

namespace {
inline int auto_lanes(double avg) {
  if (avg >= 48) return 64;
  if (avg >= 20) return 32;
  if (avg >= 10) return 16;
  return 8;
}
constexpr uint32_t kMinCut = 256;  // smaller alive sets stay on their level
}  // namespace

// ------------------------------------------------------ (b) vertex peel ---

// Removes every alive vertex of degree < k; returns the number removed.
// lvl[] of each removed vertex becomes k. A level of up to kCascadeMaxSlots
// row slots is peeled by one cascade launch (a work queue; the k-core does
// not depend on the removal order), which saves the per-round launches and
// readbacks that dominate small graphs; a larger one by synchronous rounds,
// run on the device in batches with the host reading the control block
// between batches, whose fire-and-forget decrements stream faster than the
// cascade's returning ones (set_vertex_mode forces either).
// With alive != 0 (the level's alive vertices, for a caller that restores
// the level when the k-core is empty), the peel may stop as soon as fewer
// than k + 1 vertices are left -- a non-empty k-core has at least k + 1 --
// and then returns `alive` with the level part-peeled.
uint64_t Solver::peel_vertices(LevelGraph& g, uint32_t k, uint64_t alive) {
  if (g.n == 0) return 0;
  if (alive && alive <= k) return alive;  // no k-core fits
  const uint64_t stop_at = alive ? alive - k : 0;  // removed >= stop_at: fewer than k + 1 left
  vertex_rounds_ = vertex_mode_ == 1 || (vertex_mode_ == 0 && g.slots > kCascadeMaxSlots);
  if (vertex_rounds_) return peel_vertices_rounds(g, k, stop_at, alive);
  const uint64_t cap = cascade_capacity(g.n, g.slots);
  if (cap >= 0xFFFFFFFFull) fail("vertex peel: cascade queue exceeds 2^32 items");
  vq_.ensure(sizeof(unsigned long long) * cap);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  unsigned long long* ps = nullptr;
  if (probe_kernel_ == kProbeVertexPeel) {
    if (probe_push_ & 1) {  // measurement: events at the launch's start and end
      hipEvent_t* ev = probe_events(1);
      e0 = ev[0];
      e1 = ev[1];
    }
    if (probe_push_ & 2) ps = probe_stats_.as<unsigned long long>();  // measurement: slot counters
  }
  core_cascade(g.off, g.adj, g.n, k, g.lvl, g.deg, vq_.as<unsigned long long>(), uint32_t(cap), ctl_, ps, e0, e1,
               stream_, uint32_t(std::min<uint64_t>(stop_at, 0xFFFFFFFFull)));
  const Ctl h = read_ctl();
  if (h.vq_err & (kVqOverflow | kVqStall))
    fail(h.vq_err & kVqOverflow ? "vertex peel: cascade queue overflow" : "vertex peel: cascade stalled");
  if (h.vq_err & kVqEmpty) {
    if (trace_) std::fprintf(stderr, "[kt] vertex-peel level_n=%u k=%u empty (stopped early)\n", g.n, k);
    return alive;
  }
  const uint64_t removed = uint64_t(h.vq_reserve) - h.vq_chunks;
  if (trace_) std::fprintf(stderr, "[kt] vertex-peel level_n=%u k=%u removed=%llu\n", g.n, k, (unsigned long long)removed);
  if (removed > g.n || h.vq_done != h.vq_reserve) fail("vertex peel bookkeeping mismatch");
  return removed;
}

uint64_t Solver::peel_vertices_rounds(LevelGraph& g, uint32_t k, uint64_t stop_at, uint64_t alive) {
  constexpr int kMaxBatch = 32;
  if (g.n == 0) return 0;
  q0_.ensure(sizeof(uint32_t) * (size_t(g.n) + 1));
  uint32_t* q = q0_.as<uint32_t>();
  core_round_start(ctl_, stream_);
  const HeavyList hl = heavy();
  int batch = 2;
  uint64_t queued = 0;  // every round but the last removes a vertex
  for (;;) {
    for (int i = 0; i < batch; ++i) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      unsigned long long* ps = nullptr;
      if (probe_kernel_ == kProbeVertexPeel) {
        if (probe_push_ & 1) {  // measurement: events at every push launch's start and end
          hipEvent_t* ev = probe_events(1);
          e0 = ev[0];
          e1 = ev[1];
        }
        if (probe_push_ & 2) ps = probe_stats_.as<unsigned long long>();  // measurement: slot counters
      }
      core_round(g.off, g.adj, g.n, q, ctl_, k, g.lvl, g.deg, hl, ps, e0, e1, stream_);
    }
    queued += batch;
    const Ctl h = read_ctl();
    if (trace_)
      std::fprintf(stderr, "[kt] vertex-peel level_n=%u k=%u batch=%d removed=%llu\n", g.n, k, batch, h.removed);
    if (h.action != kRoundPeel) {
      if (h.removed > g.n) fail("vertex peel bookkeeping mismatch");
      return h.removed;
    }
    if (stop_at && h.removed >= stop_at) return alive;  // fewer than k + 1 left: the k-core is empty
    if (queued > uint64_t(g.n) + 2 * kMaxBatch) fail("vertex peel did not converge");
    batch = std::min(batch * 2, kMaxBatch);
  }
}

// Induced subgraph of the alive vertices of `p` as the next level.
LevelGraph& Solver::cut_level(LevelGraph& p, uint32_t k_enter, uint32_t alive, bool recount) {
  if (levels_.size() <= depth_) levels_.emplace_back(new LevelGraph);
  LevelGraph& s = *levels_[depth_];
  // alive bitmap | word popcounts | their exclusive scan, then the ranked list
  const size_t words = (size_t(p.n) + 31) / 32;
  list_.ensure(sizeof(uint32_t) * (size_t(n_) + 1));
  const size_t cwords = (words + 31) / 32;
  new_id_.ensure(sizeof(uint32_t) * (3 * (words + 1) + cwords + 1));
  uint32_t* abits = new_id_.as<uint32_t>();
  uint32_t* acnt = abits + words + 1;
  uint32_t* apref = acnt + words + 1;
  uint32_t* acoarse = apref + words + 1;
  alive_words(p.lvl, p.n, abits, acnt, stream_);
  alive_coarse(abits, p.n, acoarse, stream_);
  exclusive_scan_u32(acnt, apref, uint32_t(words + 1));
  alive_ranked_list(abits, apref, p.n, list_.as<uint32_t>(), stream_);
  const uint32_t na = read_u32(apref + words);
  if (na != alive) fail("vertex peel bookkeeping mismatch");
  s.b_deg.ensure(sizeof(uint32_t) * (size_t(na) + 1));
  s.b_orig.ensure(sizeof(uint32_t) * (size_t(na) + 1));
  s.b_off.ensure(sizeof(uint64_t) * (size_t(na) + 1));
  s.b_lvl.ensure(sizeof(uint32_t) * (size_t(na) + 1));
  s.n = na;
  s.deg = s.b_deg.as<uint32_t>();
  s.to_orig = s.b_orig.as<uint32_t>();
  s.off = s.b_off.as<uint64_t>();
  s.lvl = s.b_lvl.as<uint32_t>();
  s.k_enter = k_enter;
  s.ok.assign(1, k_enter);
  s.exact_from = s.exact_to = 0;
  fill_u32(s.deg + na, 1, 0, stream_);
  const uint32_t* list = list_.as<uint32_t>();
  sub_degrees(list, recount ? nullptr : p.deg, p.to_orig, na, s.deg, s.to_orig, stream_);
  const uint32_t cap = uint32_t(std::min<uint64_t>(na, p.slots / 8192 + 1));
  cursor_.ensure(sizeof(uint64_t) * (size_t(cap) + 1));
  // Degrees unknown and the build's partition buffer (free after the build)
  // large enough for every kept row: fill the rows at their full lengths in
  // it while counting their alive neighbours, then pack them (one walk of
  // the kept rows instead of a counting walk and a filling walk).
  const bool one_walk = recount && part_.bytes >= sizeof(uint32_t) * std::max<uint64_t>(p.slots, 1);
  if (trace_)
    std::fprintf(stderr, "[kt] cut level_n=%u -> %u recount=%d fill=%s\n", p.n, na, int(recount),
                 one_walk ? "one-walk" : recount ? "count+fill" : "fill");
  if (one_walk) {
    keys_.ensure(sizeof(uint64_t) * (size_t(na) + 1));
    uint64_t* ub_off = keys_.as<uint64_t>();
    uint32_t* ub_adj = part_.as<uint32_t>();
    list_row_lengths(p.off, list, na, s.deg, stream_);
    exclusive_scan_u32_to_u64(s.deg, ub_off, na + 1);
    sub_fill(p.off, p.adj, list, abits, apref, acoarse, na, ub_off, ub_adj, heavy(), cursor_.as<uint64_t>(), cap,
             stream_, s.deg);
    exclusive_scan_u32_to_u64(s.deg, s.off, na + 1);
    s.slots = read_u64(s.off + na);
    s.b_adj.ensure(sizeof(uint32_t) * std::max<uint64_t>(s.slots, 1));
    s.adj = s.b_adj.as<uint32_t>();
    rows_pack(ub_adj, ub_off, s.off, s.deg, na, s.adj, stream_);
  } else {
    if (recount) list_alive_degrees(p.off, p.adj, list, na, abits, acoarse, s.deg, heavy(), stream_);
    exclusive_scan_u32_to_u64(s.deg, s.off, na + 1);
    s.slots = read_u64(s.off + na);
    s.b_adj.ensure(sizeof(uint32_t) * std::max<uint64_t>(s.slots, 1));
    s.adj = s.b_adj.as<uint32_t>();
    sub_fill(p.off, p.adj, list, abits, apref, acoarse, na, s.off, s.adj, heavy(), cursor_.as<uint64_t>(), cap, stream_);
  }
  fill_u32(s.lvl, na, kNone, stream_);
  ++depth_;
  return s;
}

uint64_t Solver::degree_at_least(uint32_t k) const {
  if (deg_at_least_.empty()) return ~uint64_t(0);
  if (k < kDegExact) return deg_at_least_[k];
  uint32_t b = kDegExact + (31u - __builtin_clz(k)) - kDegExactLog2;
  if ((uint64_t(1) << (b - kDegExact + kDegExactLog2)) < k) ++b;
  return deg_at_least_[std::min<uint32_t>(b, kDegBins)];
}

// The k-core lies inside the subgraph induced by the vertices of degree >= k
// (and is that subgraph's k-core), so it is reached by one filter of the
// vertices, a cut that counts the kept vertices' degrees from their rows,
// and a peel of the small induced subgraph, instead of a cascade through the
// whole graph.
uint64_t Solver::filtered_core(uint32_t k) {
  depth_ = 1;
  LevelGraph& g0 = *levels_[0];
  const uint32_t n = g0.n;
  degrees_from_offsets(g0.off, n, g0.deg, stream_);
  zero_u32(&ctl_->nalive);
  core_filter(g0.deg, n, k, g0.lvl, &ctl_->nalive, stream_);
  const uint32_t kept = read_u32(&ctl_->nalive);
  if (kept == 0) return 0;  // every vertex is stamped removed at k
  LevelGraph& s = cut_level(g0, k, kept, true);
  const uint64_t removed = peel_vertices(s, k);
  if (removed >= kept) {  // the k-core is empty: level 0 says so too
    depth_ = 1;
    fill_u32(g0.lvl, n, k, stream_);
    return 0;
  }
  return kept - removed;
}

// Start of the max-core search from a high threshold: a sixteenth of a lower
// bound of the degree h-index H (at least H vertices have degree >= H; the
// max core is at most H), reached by filtered_core. On power-law graphs the
// row slots of the vertices of degree >= d shrink very slowly with d, so a
// low threshold costs the cut little more than a high one, and it leaves a
// cut level that pass 2's (L-1)-core usually lies in (no second cut from
// level 0). When its core is empty the threshold drops by 4x, and when the
// filter would keep more than a quarter of the vertices the plain search
// from 1 runs instead. On success *gp is the cut level with the kf-core
// alive, *k_out = kf and *alive_out its vertex count. The degree histogram
// stays for later filtered cuts.
bool Solver::filtered_start(LevelGraph** gp, uint32_t* k_out, uint64_t* alive_out) {
  LevelGraph& g0 = **gp;
  const uint32_t n = g0.n;
  if (n < 4 * kMinCut) return false;
  dhist_.ensure(sizeof(uint32_t) * kDegBins);
  degree_hist(g0.deg, n, dhist_.as<uint32_t>(), stream_);
  // into pinned host memory (arena mode has none and copies to pageable memory)
  std::vector<uint32_t> paged(pinned_hist_ ? 0 : kDegBins);
  uint32_t* h = pinned_hist_ ? pinned_hist_ : paged.data();
  KT_HIP(hipMemcpyAsync(h, dhist_.p, sizeof(uint32_t) * kDegBins, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  // deg_at_least_[b]: vertices whose degree is at least the lower edge of bin b
  deg_at_least_.assign(kDegBins + 1, 0);
  uint64_t hidx = 0;
  for (int b = int(kDegBins) - 1; b >= 0; --b) {
    deg_at_least_[b] = deg_at_least_[b + 1] + h[b];
    const uint64_t lo = uint32_t(b) < kDegExact ? uint64_t(b) : uint64_t(1) << (b - kDegExact + kDegExactLog2);
    hidx = std::max(hidx, std::min(deg_at_least_[b], lo));
  }
  for (uint32_t kf = uint32_t(hidx / 16); kf >= 3; kf /= 4) {
    if (degree_at_least(kf) > n / 4) break;
    const uint64_t alive = filtered_core(kf);
    if (trace_)
      std::fprintf(stderr, "[kt] max-core filter k=%u (h-index >= %llu): %llu in the core\n", kf,
                   (unsigned long long)hidx, (unsigned long long)alive);
    if (alive) {
      *gp = levels_[1].get();
      *k_out = kf;
      *alive_out = alive;
      return true;
    }
  }
  // back to the whole graph
  depth_ = 1;
  degrees_from_offsets(g0.off, n, g0.deg, stream_);
  fill_u32(g0.lvl, n, kNone, stream_);
  return false;
}

// Galloping search for the largest k with a non-empty k-core. Leaves the
// max-core alive on the deepest level and returns its k (0: no edges).
uint32_t Solver::max_core(bool allow_cut, bool fresh) {
  depth_ = 1;
  LevelGraph* g = levels_.empty() || fresh ? &level0() : levels_[0].get();
  deg_bak_.ensure(sizeof(uint32_t) * (size_t(n_) + 1));
  q0_.ensure(sizeof(uint32_t) * (size_t(n_) + 1));
  q1_.ensure(sizeof(uint32_t) * (size_t(n_) + 1));
  uint64_t before = g->n;
  if (fresh) {
    degrees_from_offsets(g->off, g->n, g->deg, stream_);
    fill_u32(g->lvl, g->n, kNone, stream_);
  } else {
    zero_u32(&ctl_->nalive);
    alive_list(g->lvl, g->n, q1_.as<uint32_t>(), q0_.as<uint32_t>(), &ctl_->nalive, stream_);
    before = read_u32(&ctl_->nalive);
  }
  uint32_t k = 1, step = 16;
  uint64_t alive = 0;
  if (!(allow_cut && fresh && filtered_start(&g, &k, &alive))) {
    const uint64_t gone = peel_vertices(*g, 1);
    if (gone >= before) return 0;
    alive = before - gone;
    g->ok.assign(1, 1);
    g->exact_from = g->exact_to = 0;
    if (allow_cut && alive * 4 <= g->n && alive >= kMinCut) g = &cut_level(*g, 1, uint32_t(alive));
  }
  bool grow = true;
  uint32_t hi = ~0u;  // smallest threshold known to empty the level
  while (step > 0) {
    // A small level finishes in one workgroup that walks every threshold
    // (each costs about a row length of latency), the gallop in cascades of
    // the whole level (about log2 of the range): walk when few thresholds
    // are left, or on a cut level (its threshold already near the max core)
    // whose rows are short.
    if (vertex_mode_ == 0 && g->n <= core_decompose_max_n() && g->slots <= kDecomposeMaxSlots &&
        (uint64_t(hi) - k <= kDecomposeMaxRange ||
         (depth_ > 1 && g->slots <= uint64_t(kDecomposeMaxRow) * g->n)))
      return decompose_level(*g, k);
    const uint32_t kk = k + step;
    KT_HIP(hipMemcpyAsync(deg_bak_.p, g->deg, sizeof(uint32_t) * g->n, hipMemcpyDeviceToDevice, stream_));
    const uint64_t removed = peel_vertices(*g, kk, alive);  // an overshoot is restored below
    if (removed < alive) {
      alive -= removed;
      k = kk;
      g->ok.push_back(k);
      step = grow ? step * 2 : step / 2;
      if (allow_cut && alive * 4 <= g->n && alive >= kMinCut) g = &cut_level(*g, k, uint32_t(alive));
    } else {
      // overshoot: re-open what this step removed and restore degrees
      hi = std::min(hi, kk);
      core_reopen_above(g->lvl, g->n, k, stream_);
      KT_HIP(hipMemcpyAsync(g->deg, deg_bak_.p, sizeof(uint32_t) * g->n, hipMemcpyDeviceToDevice, stream_));
      if (step == 1) break;
      step /= 2;
      grow = false;
    }
  }
  return k;
}

// The rest of the max-core search on a small level in one workgroup
// (core_decompose): from the k-core alive on g, thresholds walk upward with
// no restores. Leaves the max core alive on g with its degrees and every
// vertex removed on the way stamped with its core number + 1, so every cut
// of g from k up to the max core is exact; returns the max core's k.
uint32_t Solver::decompose_level(LevelGraph& g, uint32_t k) {
  dec_out_.ensure(sizeof(uint32_t) * 2);
  q0_.ensure(sizeof(uint32_t) * (size_t(g.n) + 1));
  if (trace_) trace_ms();
  core_decompose(g.off, g.adj, g.n, k, g.lvl, g.deg, q0_.as<uint32_t>(), dec_out_.as<uint32_t>(), stream_);
  KT_HIP(hipMemcpyAsync(pinned_, dec_out_.p, sizeof(uint32_t) * 2, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  const uint32_t kmax = pinned_[0], alive = pinned_[1];
  if (kmax < k || alive == 0 || alive > g.n) fail("vertex peel bookkeeping mismatch (core decomposition)");
  if (trace_)
    std::fprintf(stderr, "[kt] core decomposition level_n=%u slots=%llu from k=%u: max core %u, %u vertices (%.2f ms)\n",
                 g.n, (unsigned long long)g.slots, k, kmax, alive, trace_ms());
  g.ok.push_back(kmax);
  g.exact_from = k;
  g.exact_to = kmax;
  return kmax;
}

// Leaves the k-core alive on the deepest level whose vertices all have core
// number >= its entry threshold <= k: re-open everything that level removed
// above its largest successful threshold kj <= k (k itself when the level's
// stamps are exact there), recount degrees, peel at k.
void Solver::core_at(uint32_t k) {
  size_t d = depth_ - 1;
  while (d > 0 && levels_[d]->k_enter > k) --d;
  if (d == 0 && k >= 3 && levels_[0]->n >= 4 * kMinCut && degree_at_least(k) <= levels_[0]->n / 4) {
    // no cut level holds the k-core, but few vertices have degree >= k:
    // filter and cut again rather than re-peel the whole graph
    const uint64_t alive = filtered_core(k);
    if (trace_) std::fprintf(stderr, "[kt] core_at k=%u by degree filter: %llu in the core\n", k,
                             (unsigned long long)alive);
    return;
  }
  depth_ = d + 1;
  LevelGraph& g = *levels_[d];
  uint32_t kj = std::max<uint32_t>(g.k_enter, 1);
  for (uint32_t t : g.ok)
    if (t <= k) kj = std::max(kj, t);
  const bool exact = g.exact_from && g.exact_from <= k && k <= g.exact_to;
  if (exact) kj = k;
  core_reopen_above(g.lvl, g.n, kj, stream_);
  core_alive_degrees(g.off, g.adj, g.lvl, g.n, g.deg, heavy(), stream_);
  if (k > kj) peel_vertices(g, k);
  // stamps at or below k are untouched; above k they are gone
  if (exact)
    g.exact_to = k;
  else
    g.exact_from = g.exact_to = 0;
}

// --------------------------------------------------- (a1-a3) extraction ---

// Rows of the oriented CSR in off_d_ / adj_d_ ascending: rows of up to 8192
// entries in LDS (one wave or one workgroup per row), longer rows by a
// segmented radix sort over all rows.
void Solver::sort_oriented_rows(uint32_t nc, uint32_t mc, uint32_t max_row) {
  if (max_row <= 8192) {
    sort_rows(off_d_.as<uint32_t>(), nc, adj_d_.as<uint32_t>(), max_row, stream_);
    return;
  }
  tkeys_.ensure(sizeof(uint32_t) * std::max<uint32_t>(mc, 1));
  size_t tb = 0;
  KT_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, adj_d_.as<uint32_t>(), tkeys_.as<uint32_t>(), mc,
                                                    nc, off_d_.as<uint32_t>(), off_d_.as<uint32_t>() + 1, 0, 32,
                                                    stream_));
  temp_.ensure(tb);
  KT_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(temp_.p, tb, adj_d_.as<uint32_t>(), tkeys_.as<uint32_t>(), mc,
                                                    nc, off_d_.as<uint32_t>(), off_d_.as<uint32_t>() + 1, 0, 32,
                                                    stream_));
  KT_HIP(hipMemcpyAsync(adj_d_.p, tkeys_.p, sizeof(uint32_t) * mc, hipMemcpyDeviceToDevice, stream_));
}

void Solver::extract_core(LevelGraph& g, CoreView* cv) {
  core_valid_ = false;  // the core buffers are rewritten
  const uint32_t n = g.n;
  const size_t nn = size_t(n_) + 1;
  keys_.ensure(sizeof(uint64_t) * nn);
  keys2_.ensure(sizeof(uint64_t) * nn);
  vals_.ensure(sizeof(uint32_t) * nn);
  vals2_.ensure(sizeof(uint32_t) * nn);
  to_.ensure(sizeof(uint32_t) * nn);
  zero_u32(&ctl_->nalive);
  core_collect(g.lvl, g.deg, g.to_orig, n, keys_.as<uint64_t>(), vals_.as<uint32_t>(), &ctl_->nalive, stream_);
  const uint32_t nc = read_u32(&ctl_->nalive);
  *cv = CoreView{};
  cv->nc = cv->nc0 = nc;
  inv_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  inv_orig_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  dout_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  off_d_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  sort_pairs_u64(keys_.as<uint64_t>(), keys2_.as<uint64_t>(), vals_.as<uint32_t>(), vals2_.as<uint32_t>(), nc);
  fill_u32(to_.as<uint32_t>(), n, kNone, stream_);
  core_assign_rank(keys2_.as<uint64_t>(), vals2_.as<uint32_t>(), nc, to_.as<uint32_t>(), inv_.as<uint32_t>(),
                   inv_orig_.as<uint32_t>(), stream_);
  fill_u32(dout_.as<uint32_t>(), uint64_t(nc) + 1, 0, stream_);
  orient_count(g.off, g.adj, to_.as<uint32_t>(), inv_.as<uint32_t>(), nc, dout_.as<uint32_t>(), stream_);
  exclusive_scan_u32(dout_.as<uint32_t>(), off_d_.as<uint32_t>(), nc + 1);
  const uint32_t mc = read_u32(off_d_.as<uint32_t>() + nc);
  if (mc >= 0x7FFFFFFFu) fail("core subgraph exceeds 2^31 edges");
  cv->mc = mc;
  cv->max_dout = reduce_max_u32(dout_.as<uint32_t>(), nc);
  const size_t mm = std::max<uint32_t>(mc, 1);
  adj_d_.ensure(sizeof(uint32_t) * mm);
  src_.ensure(sizeof(uint32_t) * mm);
  orient_fill(g.off, g.adj, to_.as<uint32_t>(), inv_.as<uint32_t>(), nc, off_d_.as<uint32_t>(),
              adj_d_.as<uint32_t>(), src_.as<uint32_t>(), nullptr, stream_);
  sort_oriented_rows(nc, mc, cv->max_dout);
  orig0_.ensure(sizeof(uint32_t) * mm);
  del_thr_.ensure(sizeof(uint32_t) * mm);
  sup_.ensure(sizeof(int32_t) * mm);
  state_.ensure(sizeof(uint32_t) * mm);
  eq0_.ensure(sizeof(uint32_t) * mm);
  eq1_.ensure(sizeof(uint32_t) * mm);
  iota_u32(orig0_.as<uint32_t>(), mc, stream_);
  fill_u32(del_thr_.as<uint32_t>(), mc, kNone, stream_);
  fill_u32(state_.as<uint32_t>(), mc, kNone, stream_);
  cv->inv = inv_orig_.as<uint32_t>();
  cv->off_d = cv->off_d0 = off_d_.as<uint32_t>();
  cv->adj_d = cv->adj_d0 = adj_d_.as<uint32_t>();
  cv->src = cv->src0 = src_.as<uint32_t>();
  cv->orig = orig0_.as<uint32_t>();
  cv->del_thr = del_thr_.as<uint32_t>();
  cv->mc0 = mc;
  cv->max_dout0 = cv->max_dout;
  cv->sup = sup_.as<int32_t>();
  cv->state = state_.as<uint32_t>();
  cv->work = -1;
  build_undirected_view(cv);
  if (core_mode_ == 2 ? cv->nc <= 65536 : (core_mode_ == 0 && dense_fits(cv->nc, cv->mc))) make_dense(cv);
}

// Bit-matrix rows of the core (all edges, live edges) plus per-word prefix
// popcounts of the static rows.
void Solver::make_dense(CoreView* cv) {
  const uint32_t ws = dense_words(cv->nc);
  const size_t words = size_t(cv->nc) * ws;
  bits0_.ensure(sizeof(uint64_t) * std::max<size_t>(words, 1));
  bits_.ensure(sizeof(uint64_t) * std::max<size_t>(words, 1));
  fbits_.ensure(sizeof(uint64_t) * std::max<size_t>(words, 1));
  pref0_.ensure(sizeof(uint32_t) * std::max<size_t>(words, 1));
  cv->dense = true;
  cv->ws = ws;
  cv->bits0 = bits0_.as<uint64_t>();
  cv->bits = bits_.as<uint64_t>();
  cv->fbits = fbits_.as<uint64_t>();
  cv->pref0 = pref0_.as<uint32_t>();
  dense_rows(cv->off_u, cv->adj_u, cv->nc, ws, cv->bits0, cv->bits, cv->pref0, stream_);
  fill_u64(cv->fbits, words, 0, stream_);
  cv->eidm = nullptr;
  const uint64_t eb = sizeof(uint32_t) * uint64_t(cv->nc) * cv->nc;
  if (eb && eb <= eidm_budget_) {
    eidm_.ensure(eb);
    cv->eidm = eidm_.as<uint32_t>();
    dense_eidm(cv->off_u, cv->adj_u, cv->eid_u, cv->nc, cv->eidm, stream_);
  }
}

// Undirected CSR of the current oriented edges: row w = [sources of w's
// in-edges, ascending] ++ out(w), with the oriented edge id of every slot.
void Solver::build_undirected_view(CoreView* cv) {
  const uint32_t nc = cv->nc, mc = cv->mc;
  const size_t mm = std::max<uint32_t>(mc, 1);
  din_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  din_off_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  off_u_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  tkeys_.ensure(sizeof(uint32_t) * mm);
  tvals_.ensure(sizeof(uint32_t) * mm);
  tiota_.ensure(sizeof(uint32_t) * mm);
  in_src_.ensure(sizeof(uint32_t) * mm);
  adj_u_.ensure(sizeof(uint32_t) * 2 * mm);
  eid_u_.ensure(sizeof(uint32_t) * 2 * mm);
  fill_u32(din_.as<uint32_t>(), uint64_t(nc) + 1, 0, stream_);
  count_in_degrees(cv->adj_d, mc, din_.as<uint32_t>(), stream_);
  iota_u32(tiota_.as<uint32_t>(), mc, stream_);
  if (mc) {
    int bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) < nc) ++bits;
    size_t tb = 0;
    KT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, cv->adj_d, tkeys_.as<uint32_t>(), tiota_.as<uint32_t>(),
                                              tvals_.as<uint32_t>(), mc, 0, bits, stream_));
    temp_.ensure(tb);
    KT_HIP(hipcub::DeviceRadixSort::SortPairs(temp_.p, tb, cv->adj_d, tkeys_.as<uint32_t>(), tiota_.as<uint32_t>(),
                                              tvals_.as<uint32_t>(), mc, 0, bits, stream_));
  }
  gather_u32(tvals_.as<uint32_t>(), cv->src, mc, in_src_.as<uint32_t>(), stream_);
  exclusive_scan_u32(din_.as<uint32_t>(), din_off_.as<uint32_t>(), nc + 1);
  add_offsets(din_off_.as<uint32_t>(), cv->off_d, nc + 1, off_u_.as<uint32_t>(), stream_);
  build_undirected(cv->off_d, cv->adj_d, din_off_.as<uint32_t>(), in_src_.as<uint32_t>(), tvals_.as<uint32_t>(),
                   nc, off_u_.as<uint32_t>(), adj_u_.as<uint32_t>(), eid_u_.as<uint32_t>(), stream_);
  cv->off_u = off_u_.as<uint32_t>();
  cv->adj_u = adj_u_.as<uint32_t>();
  cv->eid_u = eid_u_.as<uint32_t>();
}

// ------------------------------------------------------------ (c) count ---

void Solver::count_support(CoreView* cv, float* ms) {
  fill_u32(reinterpret_cast<uint32_t*>(cv->sup), cv->mc, 0, stream_);
  const double avg = cv->nc ? double(cv->mc) / cv->nc : 0.0;
  const int lanes = lanes_count_ ? lanes_count_ : auto_lanes(avg);
  KT_HIP(hipEventRecord(ev_[8], stream_));
  uint32_t kind = 1;
  if (cv->dense) {
    kind = 2;
    dense_support(cv->off_d, cv->adj_d, cv->nc, cv->mc, cv->bits, cv->ws, nullptr, cv->sup, stream_);
  } else if (!force_edge_count_ && support_bitmap_fits(cv->nc, cv->max_dout)) {
    kind = 0;
    support_bitmap(cv->off_d, cv->adj_d, cv->nc, cv->max_dout, lanes, cv->sup, ctl_, stream_);
  } else
    support_edges(cv->off_d, cv->adj_d, cv->src, cv->nc, cv->mc, lanes, cv->sup, stream_);
  KT_HIP(hipEventRecord(ev_[9], stream_));
  if (ms) {  // the timed count of the pass-2 core
    KT_HIP(hipEventSynchronize(ev_[9]));
    KT_HIP(hipEventElapsedTime(ms, ev_[8], ev_[9]));
    count_kind_ = kind;
    count_words_ = cv->dense ? cv->ws : 0;
  }
}

// ----------------------------------------------------- (d3) compaction ---

void Solver::compact_core(CoreView* cv, bool carry_sup, const uint32_t* vmap, uint32_t nc2) {
  const int t = cv->work == 0 ? 1 : 0;
  EdgeSet& w = work_[t];
  const uint32_t nc = cv->nc;
  dout_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  fill_u32(dout_.as<uint32_t>(), uint64_t(nc) + 1, 0, stream_);
  row_alive_count(cv->off_d, cv->state, nc, dout_.as<uint32_t>(), stream_);
  w.off.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  // offsets of the kept edges in the old row numbering
  uint32_t* roff = w.off.as<uint32_t>();
  if (vmap) {
    roff_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
    roff = roff_.as<uint32_t>();
  }
  exclusive_scan_u32(dout_.as<uint32_t>(), roff, nc + 1);
  const uint32_t mc2 = read_u32(roff + nc);
  const uint32_t max_dout = reduce_max_u32(dout_.as<uint32_t>(), nc);
  const size_t mm = std::max<uint32_t>(mc2, 1);
  w.adj.ensure(sizeof(uint32_t) * mm);
  w.src.ensure(sizeof(uint32_t) * mm);
  w.orig.ensure(sizeof(uint32_t) * mm);
  w.sup.ensure(sizeof(int32_t) * mm);
  row_compact(cv->off_d, cv->adj_d, cv->orig, cv->sup, cv->state, nc, roff, w.adj.as<uint32_t>(),
              w.src.as<uint32_t>(), w.orig.as<uint32_t>(), carry_sup ? w.sup.as<int32_t>() : nullptr, stream_);
  if (vmap) {  // rank order is kept, so rows stay sorted and u < v stays u < v
    relabel_edges(w.adj.as<uint32_t>(), w.src.as<uint32_t>(), mc2, vmap, stream_);
    relabel_offsets(roff, vmap, nc, w.off.as<uint32_t>(), stream_);
    cv->nc = nc2;
  }
  cv->off_d = w.off.as<uint32_t>();
  cv->adj_d = w.adj.as<uint32_t>();
  cv->src = w.src.as<uint32_t>();
  cv->orig = w.orig.as<uint32_t>();
  cv->sup = w.sup.as<int32_t>();
  cv->mc = mc2;
  cv->max_dout = max_dout;
  cv->work = t;
  build_undirected_view(cv);
  fill_u32(cv->state, mc2, kNone, stream_);
  if (vmap && cv->dense) make_dense(cv);
  if (!carry_sup) count_support(cv, nullptr);
}

bool Solver::try_redensify(CoreView* cv) {
  const uint32_t nc = cv->nc;
  vflag_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  vmap_.ensure(sizeof(uint32_t) * (size_t(nc) + 1));
  fill_u32(vflag_.as<uint32_t>(), uint64_t(nc) + 1, 0, stream_);
  live_vertex_flags(cv->state, cv->src, cv->adj_d, cv->mc, vflag_.as<uint32_t>(), stream_);
  exclusive_scan_u32(vflag_.as<uint32_t>(), vmap_.as<uint32_t>(), nc + 1);
  const uint32_t nc2 = read_u32(vmap_.as<uint32_t>() + nc);
  if (uint64_t(nc2) * 4 > uint64_t(nc) * 3) return false;  // rows would shrink by less than a quarter
  compact_core(cv, false, vmap_.as<uint32_t>(), nc2);
  return true;
}

// ------------------------------------------------------------- (d) peel ---

// The round rule decides on the device whether a frontier of c edges
// (supports < thr) is cheaper to drop with an exact recount of the alive
// edges than to walk triangle by triangle. Dense cores: a recount is
// alive * ws word ANDs, a frontier walk c * (ws + tri_cost * thr) random
// lookups and atomics, so a few percent of the alive edges already favour
// the recount. Sorted-row cores pay a compaction and a rebuild first: the
// reference's third (src/kc.cu:988) scaled by drop_ratio.
// Truss levels on the core view, upward from threshold thr0 (an edge stays
// while its support >= thr, i.e. in the (thr+2)-truss). Empty levels are
// skipped by jumping to the minimum live support. With early_ans != 0, a
// first level of exactly early_ans edges at kmax early_k ends the search
// (src/kc.cu:1240-1241).
RoundRule Solver::round_rule(const CoreView& cv) const {
  RoundRule rule{};
  rule.dense = cv.dense ? 1u : 0u;
  rule.ws = cv.ws;
  rule.mc = cv.mc;
  rule.tri_cost = dense_tri_cost_;
  rule.drop_fixed = dense_drop_fixed_;
  rule.drop_ratio = sparse_drop_ratio_;
  return rule;
}

PeelArgs Solver::peel_args(const CoreView& cv) {
  PeelArgs a{};
  a.q0 = eq0_.as<uint32_t>();
  a.q1 = eq1_.as<uint32_t>();
  a.src = cv.src;
  a.dst = cv.adj_d;
  a.orig = cv.orig;
  a.state = cv.state;
  a.del_thr = cv.del_thr;
  a.sup = cv.sup;
  a.mc = cv.mc;
  a.nc = cv.nc;
  a.ws = cv.ws;
  a.off_u = cv.off_u;
  a.adj_u = cv.adj_u;
  a.eid_u = cv.eid_u;
  a.off_d = cv.off_d;
  if (cv.dense) {
    a.bits = cv.bits;
    a.fbits = cv.fbits;
    a.bits0 = cv.bits0;
    a.pref0 = cv.pref0;
    a.eidm = cv.eidm;
  }
  return a;
}

void Solver::start_rounds(CoreView* cv, int32_t thr, uint32_t r, uint64_t alive, bool reset,
                          const SearchRule& search) {
  const PeelArgs a = peel_args(*cv);
  zero_u32(&ctl_->qcount[0]);
  edge_frontier_init(cv->state, cv->sup, cv->mc, thr, a.q0, &ctl_->qcount[0], stream_);
  const RoundRule rule = round_rule(*cv);
  round_start(ctl_, r, alive, thr, rule, reset, search, stream_);
  round_stamp(a, ctl_, rule, stream_);
}

// Truss levels on the core view, upward from threshold thr0 (an edge stays
// while its support >= thr, i.e. in the (thr+2)-truss). Each completed level
// moves the threshold to its minimum live support + 1, so empty levels are
// skipped. With early_ans != 0, a first level of exactly early_ans edges at
// kmax early_k (any kmax when exact_early) ends the search (src/kc.cu:1240-1241).
// Rounds and level changes run on the device; the host reads the control
// block between batches and handles dropped frontiers and compactions.
TrussOutcome Solver::truss_search(CoreView* cv, int32_t thr0, uint64_t early_ans, uint32_t early_k,
                                  bool exact_early) {
  if (cv->mc == 0) return TrussOutcome{};
  SearchRule search{};
  search.early_ans = early_ans;
  search.early_k = early_k;
  search.exact_early = exact_early ? 1u : 0u;
  uint32_t r = 1;
  return run_levels(cv, thr0, cv->mc, &r, search, true);
}

// Runs levels from threshold thr with `alive` live edges on the device until
// the search rule ends it (or the level empties); handles dropped frontiers
// and compactions. *r is the next round stamp on entry and on exit.
TrussOutcome Solver::run_levels(CoreView* cv, int32_t thr, uint64_t alive, uint32_t* r_io, const SearchRule& search,
                                bool allow_redensify) {
  constexpr int kMaxBatch = 32;
  TrussOutcome out;
  uint32_t r = *r_io;
  start_rounds(cv, thr, r, alive, true, search);
  int batch = 2;
  uint64_t queued = 0;  // rounds since the last restart; each removes an edge or changes level
  for (;;) {
    if (queued > 3 * uint64_t(cv->mc) + 4 * kMaxBatch) fail("edge peel did not converge");
    const double avg = cv->nc ? double(2ull * cv->mc) / cv->nc : 0.0;
    LaunchProbe probe;
    if (probe_kernel_ == kProbeDensePeel && cv->dense) {  // measurement of the dense peel round
      if (probe_push_ & 1) probe.ev = probe_events(batch);
      if (probe_push_ & 2) probe.stats = probe_stats_.as<unsigned long long>();
    }
    peel_rounds(peel_args(*cv), ctl_, round_rule(*cv), lanes_peel_ ? lanes_peel_ : auto_lanes(avg), batch, stream_,
                probe);
    queued += batch;
    const Ctl h = read_ctl();
    if (trace_)
      std::fprintf(stderr, "[kt] edge-peel batch=%d action=%u thr=%d alive=%llu mc=%u frontier=%u levels=%u\n",
                   batch, h.action, h.thr, h.alive, cv->mc, h.qcount[h.cur], h.levels);
    if (h.err & kErrPeelCount) fail("edge peel bookkeeping mismatch (live count)");
    out.rounds = h.rounds_done;
    out.levels = h.levels;
    out.kmax = h.level_kmax;
    out.count = h.level_count;
    out.floor_thr = h.level_floor;
    if (h.action == kRoundPeel || h.action == kRoundMinSup || h.action == kRoundSweep) {
      batch = std::min(batch * 2, kMaxBatch);
      continue;
    }
    out.alive = h.alive;
    r = h.round + 1;
    if (h.action == kRoundDone || h.action == kRoundEmpty) {
      out.cut_short = h.action == kRoundDone && h.alive > 0 && search.max_levels && h.levels >= search.max_levels;
      break;
    }
    alive = h.alive;
    thr = h.thr;
    if (h.action == kRoundDrop) {  // drop the stamped frontier, recount the rest exactly
      ++out.recounts;
      if (cv->dense) {
        if (allow_redensify && try_redensify(cv)) {  // smaller matrix on the vertices left, exact recount
          if (cv->mc != alive) fail("edge peel bookkeeping mismatch (redensify)");
          ++out.compactions;
          r = 1;
        } else {
          frontier_clear(peel_args(*cv), ctl_, stream_);
          dense_support(cv->off_d, cv->adj_d, cv->nc, cv->mc, cv->bits, cv->ws, cv->state, cv->sup, stream_);
        }
      } else {
        compact_core(cv, false);
        ++out.compactions;
        alive = cv->mc;
        r = 1;
      }
    } else if (h.action == kRoundCompact) {  // half the edges are dead: compact, keep supports
      compact_core(cv, true);
      ++out.compactions;
      if (cv->mc != alive) fail("edge peel bookkeeping mismatch (compaction)");
      r = 1;
    } else {
      fail("edge peel: unexpected round action");
    }
    start_rounds(cv, thr, r, alive, false, search);
    batch = 2;
    queued = 0;
  }
  *r_io = r;
  return out;
}

// The largest threshold t whose level is non-empty, on a dense view. The
// first walk_levels_ levels are walked as truss_search does; if the search is
// not over by then, thresholds are probed instead: each probe peels one
// level at t from the last non-empty state (kept in a backup) and is undone
// when it empties the view. Thresholds gallop up by doubling steps from the
// minimum support (never past empty_from - 1, a bound known to hold, so a
// view whose answer sits at the bound is done by one kept probe), then
// bisect until at most kWalkWindow thresholds are left, whose levels are
// then walked. The result equals the level walk's: the last
// non-empty probe leaves the t-level, whose minimum support is t* (a level at
// t* + 1 is empty), so kmax = t* + 2 with its edge count, and every edge
// deleted on the way carries a threshold <= the floor. (The reference's
// speculative k-steps with backups, src/kc.cu:1189-1196, 1252-1272.)
TrussOutcome Solver::truss_gallop(CoreView* cv, int64_t empty_from) {
  const int64_t kWalkWindow = walk_window_;  // thresholds left when probing gives way to walking
  TrussOutcome out;
  if (cv->mc == 0) return out;
  SearchRule walk{};
  walk.max_levels = walk_levels_;
  uint32_t r = 1;
  if (trace_) trace_ms();
  TrussOutcome lv = run_levels(cv, 1, cv->mc, &r, walk, true);
  if (trace_)
    std::fprintf(stderr, "[kt] gallop walk of %u levels: lo=%d alive=%llu rounds=%llu (%.2f ms)\n", lv.levels,
                 int(lv.kmax) - 2, (unsigned long long)lv.count, (unsigned long long)lv.rounds, trace_ms());
  if (!lv.cut_short) return lv;  // the walk finished the search
  out.rounds += lv.rounds;
  out.levels += lv.levels;
  out.compactions += lv.compactions;
  out.recounts += lv.recounts;
  out.kmax = lv.kmax;
  out.count = lv.count;
  out.floor_thr = lv.floor_thr;
  const size_t mc = cv->mc, words = size_t(cv->nc) * cv->ws;
  bk_state_.ensure(sizeof(uint32_t) * std::max<size_t>(mc, 1));
  bk_sup_.ensure(sizeof(int32_t) * std::max<size_t>(mc, 1));
  bk_bits_.ensure(sizeof(uint64_t) * std::max<size_t>(words, 1));
  bk_del_.ensure(sizeof(uint32_t) * std::max<size_t>(cv->mc0, 1));
  auto copy = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) KT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_));
  };
  auto save = [&]() {
    copy(bk_state_.p, cv->state, sizeof(uint32_t) * mc);
    copy(bk_sup_.p, cv->sup, sizeof(int32_t) * mc);
    copy(bk_bits_.p, cv->bits, sizeof(uint64_t) * words);
    copy(bk_del_.p, cv->del_thr, sizeof(uint32_t) * cv->mc0);
  };
  auto restore = [&]() {
    copy(cv->state, bk_state_.p, sizeof(uint32_t) * mc);
    copy(cv->sup, bk_sup_.p, sizeof(int32_t) * mc);
    copy(cv->bits, bk_bits_.p, sizeof(uint64_t) * words);
    copy(cv->del_thr, bk_del_.p, sizeof(uint32_t) * cv->mc0);
  };
  SearchRule single{};
  single.max_levels = 1;
  int64_t lo = int64_t(out.kmax) - 2;  // non-empty: the current state is its level
  int64_t hi = empty_from;              // smallest threshold known to be empty (-1: none)
  bool probed_empty = false;            // hi comes from a probe, not from the bound
  uint64_t alive = out.count;
  int64_t step = 1;
  while (hi < 0 || hi - lo > 1) {
    if (probed_empty && hi - lo <= kWalkWindow) {
      // A short window is left: walk its levels. A probe that empties the
      // view costs a full cascade and is then undone; the walk's only empty
      // level is the last one.
      SearchRule all{};
      const TrussOutcome w = run_levels(cv, int32_t(lo + 1), alive, &r, all, true);
      out.rounds += w.rounds;
      out.levels += w.levels;
      out.compactions += w.compactions;
      out.recounts += w.recounts;
      if (w.kmax >= 2 && w.count > 0) {
        out.kmax = w.kmax;
        out.count = w.count;
        out.floor_thr = w.floor_thr;
      }
      if (trace_)
        std::fprintf(stderr, "[kt] gallop walk from t=%lld -> kmax=%u count=%llu levels=%u rounds=%llu (%.2f ms)\n",
                     (long long)(lo + 1), out.kmax, (unsigned long long)out.count, w.levels,
                     (unsigned long long)w.rounds, trace_ms());
      break;
    }
    // gallop up by doubling steps, never past the bound; bisect once a
    // probe has emptied the view
    const int64_t t = hi < 0 ? lo + step : !probed_empty ? std::min(lo + step, hi - 1) : lo + (hi - lo) / 2;
    save();
    const TrussOutcome p = run_levels(cv, int32_t(t), alive, &r, single, false);
    out.rounds += p.rounds;
    out.levels += 1;
    out.compactions += p.compactions;
    out.recounts += p.recounts;
    if (p.alive > 0 && p.kmax >= 2) {
      lo = int64_t(p.kmax) - 2;  // >= t
      alive = p.alive;
      out.kmax = p.kmax;
      out.count = p.count;
      out.floor_thr = p.floor_thr;
      step *= gallop_growth_;
    } else {
      restore();
      hi = t;
      probed_empty = true;
    }
    if (trace_)
      std::fprintf(stderr, "[kt] gallop probe t=%lld -> %s lo=%lld hi=%lld alive=%llu rounds=%llu recounts=%u (%.2f ms)\n",
                   (long long)t, p.alive ? "kept" : "empty", (long long)lo, (long long)hi, (unsigned long long)alive,
                   (unsigned long long)p.rounds, p.recounts, trace_ms());
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

uint32_t Solver::decompose_pairs_device(const uint32_t* d_pairs, uint64_t m, uint32_t n, CoreView* out) {
  KT_HIP(hipSetDevice(device_));
  have_truss_ = false;
  build_from_pairs(d_pairs, m, n);
  depth_ = 1;
  LevelGraph& g = level0();
  degrees_from_offsets(g.off, g.n, g.deg, stream_);
  fill_u32(g.lvl, g.n, kNone, stream_);
  if (g.n) peel_vertices(g, 2);  // a vertex of degree < 2 is in no triangle
  CoreView cv;
  extract_core(g, &cv);
  uint32_t kmax = 0;
  if (cv.mc) {
    count_support(&cv, nullptr);
    kmax = truss_search(&cv, 1, 0, 0).kmax;
  }
  KT_HIP(hipStreamSynchronize(stream_));
  *out = cv;
  return kmax;
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
  res->n = n_;
  res->m = m_;
  have_truss_ = false;
  last_core_ = CoreView{};
  last_truss_ = 0;
  last_floor_ = -1;
  KT_HIP(hipEventRecord(ev_[1], stream_));

  res->max_core = max_core();
  KT_HIP(hipEventRecord(ev_[2], stream_));
  if (res->max_core == 0) {  // no edges
    KT_HIP(hipEventSynchronize(ev_[2]));
    KT_HIP(hipEventElapsedTime(&res->ms_total, ev_[0], ev_[2]));
    KT_HIP(hipEventElapsedTime(&res->ms_build, ev_[0], ev_[1]));
    KT_HIP(hipEventElapsedTime(&res->ms_kcore, ev_[1], ev_[2]));
    have_truss_ = true;
    core_valid_ = true;  // empty view
    return;
  }

  // pass 1 on the max-core
  CoreView cv;
  extract_core(*levels_[depth_ - 1], &cv);
  count_support(&cv, nullptr);
  // the max-core's kmax is at most max_core + 1 (a k-truss is a
  // (k-1)-core), so its threshold max_core is empty
  TrussOutcome p1 =
      cv.dense && gallop_ ? truss_gallop(&cv, int64_t(res->max_core)) : truss_search(&cv, 1, 0, 0);
  res->k_hat = p1.kmax;
  res->k_hat_edges = p1.count;
  KT_HIP(hipEventRecord(ev_[3], stream_));

  // pass 2 on the (L-1)-core
  const uint32_t L = std::max<uint32_t>(p1.kmax, 3);
  res->core_k = L - 1;
  core_at(L - 1);
  KT_HIP(hipEventRecord(ev_[4], stream_));
  extract_core(*levels_[depth_ - 1], &cv);
  res->n_core = cv.nc;
  res->m_core = cv.mc;
  KT_HIP(hipEventRecord(ev_[5], stream_));
  count_support(&cv, &res->ms_count_kernel);
  res->lanes = lanes_count_ ? lanes_count_ : auto_lanes(cv.nc ? double(cv.mc) / cv.nc : 0.0);
  KT_HIP(hipEventRecord(ev_[6], stream_));
  TrussOutcome p2 = truss_search(&cv, int32_t(L) - 2, p1.kmax >= 3 ? p1.count : 0, p1.kmax);
  KT_HIP(hipEventRecord(ev_[7], stream_));
  KT_HIP(hipEventSynchronize(ev_[7]));
  res->kmax = p2.kmax;
  res->truss_edges = p2.kmax ? p2.count : 0;
  res->peel_rounds = p1.rounds + p2.rounds;
  res->compactions = p1.compactions + p2.compactions;
  res->recounts = p1.recounts + p2.recounts;
  res->levels = p1.levels + p2.levels;
  float t_a = 0, t_b = 0;
  KT_HIP(hipEventElapsedTime(&res->ms_total, ev_[0], ev_[7]));
  KT_HIP(hipEventElapsedTime(&res->ms_build, ev_[0], ev_[1]));
  KT_HIP(hipEventElapsedTime(&t_a, ev_[1], ev_[2]));
  KT_HIP(hipEventElapsedTime(&res->ms_pass1, ev_[2], ev_[3]));
  KT_HIP(hipEventElapsedTime(&t_b, ev_[3], ev_[4]));
  res->ms_kcore = t_a + t_b;
  KT_HIP(hipEventElapsedTime(&res->ms_extract, ev_[4], ev_[5]));
  KT_HIP(hipEventElapsedTime(&res->ms_count, ev_[5], ev_[6]));
  KT_HIP(hipEventElapsedTime(&res->ms_peel, ev_[6], ev_[7]));
  last_core_ = cv;
  core_valid_ = true;
  last_floor_ = p2.floor_thr;
  last_truss_ = res->truss_edges;
  have_truss_ = true;
}

void Solver::last_stats(kt_stats* out) {
  if (!have_truss_) fail("no solve has completed on this context");
  if (!core_valid_) fail("the last solve's core view was replaced by a trussness computation");
  std::memset(out, 0, sizeof(*out));
  const CoreView& cv = last_core_;
  out->n_core = cv.nc0;
  out->count_kernel = count_kind_;
  out->core_words = count_words_;
  out->vertex_rounds = vertex_rounds_ ? 1u : 0u;
  out->m_core = cv.mc0;
  out->truss_edges = last_truss_;
  if (cv.mc0 == 0) return;
  out->max_out_degree = cv.max_dout0;
  DevBuf tmp;
  tmp.ensure(16);
  fill_u64(tmp.as<uint64_t>(), 2, 0, stream_);
  const int32_t floor = last_truss_ ? last_floor_ : INT_MAX;
  w_stats(cv.off_d0, cv.src0, cv.adj_d0, cv.del_thr, cv.mc0, floor, tmp.as<uint64_t>(), stream_);
  uint64_t h[2];
  KT_HIP(hipMemcpyAsync(h, tmp.p, 16, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  tmp.release();
  out->w_core = h[0];
  out->w_peeled = h[1];
}

// The kmax-truss of the last solve as (u, v) pairs in input ids, written
// once into truss_out_ (from the core view's deletion thresholds).
void Solver::materialize_truss() {
  if (truss_ready_) return;
  if (!have_truss_) fail("no solve has completed on this context");
  if (!core_valid_) fail("the last solve's core view is gone");
  const CoreView& cv = last_core_;
  truss_count_ = 0;
  if (cv.mc0 && last_truss_) {
    truss_out_.ensure(sizeof(uint32_t) * 2 * size_t(cv.mc0));
    zero_u32(&ctl_->nalive);
    truss_pairs(cv.del_thr, cv.mc0, last_floor_, cv.src0, cv.adj_d0, cv.inv, truss_out_.as<uint32_t>(),
                &ctl_->nalive, stream_);
    truss_count_ = read_u32(&ctl_->nalive);
  }
  truss_ready_ = true;
}

uint64_t Solver::trussness(uint32_t* host_triples, uint64_t cap) {
  KT_HIP(hipSetDevice(device_));
  if (!n_) return 0;
  if (!triples_ready_) {
    // full decomposition of the last input's CSR (level 0 is never modified
    // after its build); the core view buffers are reused, so the last
    // solve's kmax-truss is saved first
    if (have_truss_ && core_valid_) materialize_truss();
    core_valid_ = false;
    depth_ = 1;
    LevelGraph& g = level0();
    degrees_from_offsets(g.off, g.n, g.deg, stream_);
    fill_u32(g.lvl, g.n, kNone, stream_);
    peel_vertices(g, 2);  // a vertex of degree < 2 is in no triangle
    CoreView cv;
    extract_core(g, &cv);
    triples_count_ = 0;
    if (cv.mc0) {
      count_support(&cv, nullptr);
      const uint32_t kall = truss_search(&cv, 1, 0, 0).kmax;
      out_pairs_.ensure(sizeof(uint32_t) * 3 * size_t(cv.mc0));
      zero_u32(&ctl_->nalive);
      truss_triples(cv, kall, out_pairs_.as<uint32_t>(), &ctl_->nalive, stream_);
      triples_count_ = read_u32(&ctl_->nalive);
    }
    triples_ready_ = true;
  }
  const uint64_t take = std::min<uint64_t>(triples_count_, cap);
  if (take && host_triples)
    KT_HIP(hipMemcpyAsync(host_triples, out_pairs_.p, sizeof(uint32_t) * 3 * take, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  return triples_count_;
}

uint64_t Solver::truss_edges_to_device(uint32_t* dev_pairs, uint64_t cap) {
  KT_HIP(hipSetDevice(device_));
  materialize_truss();
  const uint64_t take = std::min<uint64_t>(truss_count_, cap);
  if (take && dev_pairs)
    KT_HIP(hipMemcpyAsync(dev_pairs, truss_out_.p, sizeof(uint32_t) * 2 * take, hipMemcpyDeviceToDevice, stream_));
  return truss_count_;
}

uint64_t Solver::truss_edges(uint32_t* host_pairs, uint64_t cap) {
  KT_HIP(hipSetDevice(device_));
  materialize_truss();
  const uint64_t take = std::min<uint64_t>(truss_count_, cap);
  if (take && host_pairs)
    KT_HIP(hipMemcpyAsync(host_pairs, truss_out_.p, sizeof(uint32_t) * 2 * take, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  return truss_count_;
}

}  // namespace kt
