// Kernels of the batched solve as one disjoint union (kt_batch_union.hip).
#pragma once
#include "kt_solver.h"

namespace kt {

// status[g] |= 1 for a graph with a pair that is not canonical (u < v < n) or
// not strictly above its predecessor; status must start at 0.
void batch_validate(const uint32_t* pairs, const uint64_t* poff, const uint32_t* ns, uint32_t count,
                    uint64_t total, uint32_t* status, hipStream_t s);
// Pairs of the graphs with status 0, shifted by vstart[g], from position upoff[g].
void batch_union(const uint32_t* pairs, const uint64_t* poff, uint32_t count, uint64_t total, const uint32_t* vstart,
                 const uint32_t* status, const uint64_t* upoff, uint32_t* out, hipStream_t s);
// Per graph (vstart: count + 1 union vertex starts) the largest truss number
// of its core edges (gk, start at 0) and how many edges have it (gcnt, start
// at 0); keep (nullable, start at 0) marks those edges' input pairs.
void batch_truss(const CoreView& cv, uint32_t kall, const uint32_t* vstart, uint32_t count, uint32_t* gk,
                 unsigned long long* gcnt, const uint32_t* pairs, const uint64_t* poff, uint8_t* keep,
                 hipStream_t s);
void batch_results(const uint32_t* gk, const unsigned long long* gcnt, const uint32_t* status, uint32_t count,
                   uint32_t* kmax, unsigned long long* edges, uint32_t* status_out, hipStream_t s);
// (u, v, t) triples (u < v, input ids) of the core edges of a decomposition
// with truss number t >= 3, unordered; count (starts at 0) gets how many.
void truss_triples(const CoreView& cv, uint32_t kall, uint32_t* out, uint32_t* count, hipStream_t s);
void fill_u8(uint8_t* p, uint64_t n, uint8_t v, hipStream_t s);

}  // namespace kt
