// Shared definitions for the MI355X k-max-truss library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace kt {

constexpr uint32_t kNone = 0xFFFFFFFFu;  // "alive" stamp / "not in core" rank
constexpr int kWave = 64;                // CDNA wavefront width
constexpr int kBlock = 256;              // default workgroup: 4 waves

struct Error {
  hipError_t code;
  std::string what;
};

#define KT_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::kt::Error{e_, std::string(#expr) + " -> " + hipGetErrorString(e_) +    \
                                " at " + __FILE__ + ":" + std::to_string(__LINE__)}; \
  } while (0)

inline void fail(const std::string& what) { throw Error{hipErrorInvalidValue, what}; }

// Scratch for rows longer than one wave should walk: a device list of row
// ids and its counter; capacity = the graph's vertex count.
struct HeavyList {
  uint32_t* items;
  uint32_t* count;
  uint32_t* slices;  // capacity + 1 entries
};

// Device-side control block: counters used by the queue-driven kernels.
// Lives in device memory, one per solver context.
struct Ctl {
  uint32_t qcount[2];    // frontier queue lengths (ping-pong)
  uint32_t err;          // input validation flags
  uint32_t work;         // dynamic work-grab counter
  uint32_t nalive;       // scratch counter (compaction)
  uint32_t minsup;       // min support among live edges
  uint32_t pad[2];
  // Device-driven peel rounds. The kernels of a round read which queue holds
  // the frontier (cur), what to do with it (action) and the round stamp from
  // here, and the last workgroup of a round's closing kernel (ticket) decides
  // the next round, so the host can enqueue many rounds between readbacks.
  uint32_t cur;
  uint32_t action;       // kRound* below
  uint32_t round;        // stamp of the current frontier
  uint32_t rounds_done;  // frontiers stamped
  uint32_t ticket;       // last-workgroup detection
  uint32_t stamped;      // the current frontier has been stamped
  unsigned long long alive;    // live, unstamped edges / alive vertices
  unsigned long long removed;  // vertices removed by the vertex peel
  // Truss levels, also advanced on the device: the threshold of the level
  // being peeled (an edge stays while its support >= thr) and the record of
  // the last completed level.
  int32_t thr;
  uint32_t levels;
  uint32_t first;        // no level completed yet
  uint32_t level_kmax;   // min support of the last completed level + 2
  unsigned long long level_count;
  int32_t level_floor;   // threshold of the last completed level
  uint32_t early_k;      // early exit: a first level of early_ans edges ...
  unsigned long long early_ans;
  uint32_t exact_early;  // ... at kmax early_k (or any kmax when set)
  uint32_t max_levels;   // stop after this many completed levels (0: no limit)
  // Vertex-peel cascade queue (one launch per threshold): items reserved,
  // claimed and finished, chunk items added for long rows, and an error
  // flag that makes every wave leave.
  uint32_t vq_reserve;
  uint32_t vq_head;
  uint32_t vq_done;
  uint32_t vq_chunks;
  uint32_t vq_err;
  uint32_t vq_removed;   // vertices queued so far (a lower bound while waves run)
};

// ctl->vq_err values
constexpr uint32_t kVqOverflow = 1;  // an item past the queue's capacity
constexpr uint32_t kVqStall = 2;     // a wave waited past its spin limit
constexpr uint32_t kVqEmpty = 8;     // vertex cascade: too few vertices left for a k-core (not an error)

// ctl->err bit: a level ended with live edges but no live support (or the
// reverse); the edge-peel bookkeeping is inconsistent
constexpr uint32_t kErrPeelCount = 1u << 8;

constexpr uint32_t kRoundIdle = 0;     // vertex peel: nothing left to remove
constexpr uint32_t kRoundPeel = 1;     // walk the frontier's triangles
constexpr uint32_t kRoundDrop = 2;     // drop the frontier, recount the rest
constexpr uint32_t kRoundEmpty = 3;    // the level empties: every live edge dies now
constexpr uint32_t kRoundCompact = 4;  // compact the edge set first (unstamped)
constexpr uint32_t kRoundMinSup = 5;   // level done: reduce the live supports
constexpr uint32_t kRoundSweep = 6;    // next level: collect its frontier
constexpr uint32_t kRoundDone = 7;     // search finished (early exit, no edges)

// Measurement of a round kernel (all optional): ev holds a start and an end
// event per launch, stamped at the kernel's own start and end; stats gains
// [0] the oriented out-degrees d+(u) + d+(v) of the frontier edges walked
// and [1] their count.
struct LaunchProbe {
  hipEvent_t* ev = nullptr;
  unsigned long long* stats = nullptr;
};

// Host-set parameters of the device round decision (drop or walk a frontier).
struct RoundRule {
  uint32_t dense;     // bit-matrix core
  uint32_t ws;        // words per bit-matrix row
  uint32_t mc;        // edges of the current edge set (compaction rule)
  uint32_t pad;
  double tri_cost;    // dense: cost of one triangle visit in word ANDs
  double drop_ratio;  // sparse: drop when frontier * ratio >= alive
  double drop_fixed;  // dense: fixed cost of a drop (host round trip, restart) in word ANDs
};

// How a truss search ends: early exit on a first level of early_ans edges at
// kmax early_k (any kmax with exact_early), or after max_levels completed
// levels (0: no limit).
struct SearchRule {
  unsigned long long early_ans;
  uint32_t early_k;
  uint32_t exact_early;
  uint32_t max_levels;
  uint32_t pad;
};

// Arrays of a core view that the edge-peel round kernels touch.
struct PeelArgs {
  uint32_t *q0, *q1;
  const uint32_t *src, *dst, *orig;  // oriented edges and their base ids
  uint32_t *state, *del_thr;
  int32_t* sup;
  uint32_t mc, nc, ws;
  const uint32_t *off_u, *adj_u, *eid_u;  // undirected view (sorted rows)
  const uint32_t* off_d;                  // oriented rows (out-degrees, for measurement)
  uint64_t *bits, *fbits;                 // dense views: live / frontier bits
  const uint64_t* bits0;
  const uint32_t *pref0, *eidm;
};

}  // namespace kt
