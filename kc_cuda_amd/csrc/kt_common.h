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
  // device-driven peel rounds: which queue is current, what this round does,
  // the round stamp, and running totals
  uint32_t cur;
  uint32_t action;       // kRound* below
  uint32_t round;
  uint32_t rounds_done;
  unsigned long long alive;    // live edges (edge peel)
  unsigned long long removed;  // vertices removed (vertex peel)
};

constexpr uint32_t kRoundIdle = 0;   // empty frontier: nothing to do
constexpr uint32_t kRoundPeel = 1;   // walk the frontier's triangles
constexpr uint32_t kRoundDrop = 2;   // drop the frontier, recount the rest
constexpr uint32_t kRoundEmpty = 3;  // the frontier is every live edge

}  // namespace kt
