// Batched solve of many independent graphs (device side and its launcher).
#pragma once
#include "kt_common.h"

namespace kt {

constexpr int kBatchPhases = 10;

struct BatchArgs {
  const uint32_t* pairs;      // all graphs' canonical pairs, interleaved
  const uint64_t* poff;       // count + 1 pair offsets
  const uint32_t* ns;         // vertex count per graph
  const uint32_t* order;      // graph ids in processing order
  const uint8_t* lanes;       // lanes per edge for each graph (4..64), null: 16
  uint32_t G;                 // entries of order
  uint32_t* kmax;             // per graph
  unsigned long long* count;  // per graph: edges in the kmax-truss
  uint32_t* status;           // per graph: 0 ok, 1 invalid input
  uint8_t* keep;              // per pair: 1 when in the kmax-truss (nullable)
  unsigned long long* ticks;  // per graph solve time in wall-clock ticks (nullable)
  unsigned long long* phases; // per graph kBatchPhases entries (nullable): ticks at the end of
                              // build, cores, count 1, peel 1, count 2, peel 2; then
                              // levels << 32 | rounds of pass 1 and of pass 2; the
                              // absolute start and end ticks
  uint32_t* slab;             // grid * slab_words scratch words
  uint64_t slab_words;
  uint32_t* work;             // work-stealing counter
};

int batch_block();
uint64_t batch_slab_words(uint32_t max_n, uint64_t max_m);
// resets the work counter and launches the persistent solve
void batch_launch(const BatchArgs& a, int grid, hipStream_t s);

}  // namespace kt
