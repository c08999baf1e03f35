// Allocation-free solve over a fixed per-device scratch array (kt_noalloc.hip).
#pragma once
#include <cstdint>

#include "../../include/ktruss.h"

namespace kt {

// Edge-id matrix cap of a dense core in the allocation-free solve.
constexpr uint64_t kNoAllocEidmBytes = uint64_t(32) << 20;

// Upper bound of the scratch bytes a solve of m edges on n vertices carves
// (the serving footprint with the allocation-free solve's matrix cap).
uint64_t noalloc_footprint(uint64_t m, uint32_t n);

// Solves m device-resident canonical pairs with every buffer carved from the
// device's scratch array; writes at most cap kmax-truss pairs to d_truss
// (when non-null) and returns the truss size. Throws kt::Error when the graph
// exceeds the scratch bound or the scratch runs out.
uint64_t noalloc_solve_pairs(int device, const uint32_t* d_pairs, uint64_t m, uint32_t n, uint32_t* d_truss,
                             uint64_t cap, kt_result* res);

}  // namespace kt
