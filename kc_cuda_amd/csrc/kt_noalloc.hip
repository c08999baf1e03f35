// Allocation-free solve (kt_solve_pairs_device_noalloc in include/ktruss.h):
// the same device-resident two-pass solve as kt_solve_pairs_device, for
// callers that cannot allocate device memory. Every buffer the solve needs
// -- CSR, vertex-peel levels, core views, supports, queues, sort and scan
// temporaries, the control block -- is carved from `scratch`, a file-scope
// device array of kNoAllocScratchBytes per GPU, held from the library's load
// until the process ends. No hipMalloc, hipFree or allocating library call is
// made by a solve; the first call on a device sets up that device's solver
// object (its stream and timing events) and looks up the array's address.
//
// Every call on one device shares that device's array, which keeps the last
// call's contents: the control block is cleared at the start of each call,
// and every other buffer is written by the solve before it is read. Calls on
// one device are serialised by a host mutex; calls on one device from several
// processes overlap on the same array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/ktruss.h"
#include "kt_kernels.h"
#include "kt_noalloc.h"
#include "kt_solver.h"

static_assert(KT_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must be a multiple of 256 bytes");

static __device__ __align__(256) uint8_t scratch[KT_NOALLOC_SCRATCH_BYTES];

namespace kt {

namespace {
constexpr int kMaxDevices = 64;

struct NoAllocDevice {
  std::mutex mu;
  uint8_t* ptr = nullptr;    // this device's copy of `scratch` (device address)
  Arena arena;
  Solver* solver = nullptr;  // lives until the process ends
};

NoAllocDevice g_dev[kMaxDevices];

// Once per device (its mutex held), with that device current: the device
// address of its copy of the array (the host-side symbol is not a device
// address) and the solver object that carves from it.
void set_up(NoAllocDevice& d, int device) {
  if (!d.solver) {
    KT_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(std::string("kt_solve_pairs_device_noalloc: hipGetSymbolAddress(scratch) failed: ") +
           hipGetErrorString(e));
    d.ptr = static_cast<uint8_t*>(p);
    d.arena.base = d.ptr;
    d.arena.size = KT_NOALLOC_SCRATCH_BYTES;
    Solver* s = new Solver(device, &d.arena);
    s->set_eidm_budget(kNoAllocEidmBytes);
    d.solver = s;
  }
}
}  // namespace

uint64_t noalloc_footprint(uint64_t m, uint32_t n) {
  return 160ull * m + 64ull * uint64_t(n) + std::min<uint64_t>(72ull * m, kNoAllocEidmBytes) +
         std::min<uint64_t>(512ull * m, kNoAllocEidmBytes) + (16ull << 20);
}

uint64_t noalloc_solve_pairs(int device, const uint32_t* d_pairs, uint64_t m, uint32_t n, uint32_t* d_truss,
                             uint64_t cap, kt_result* res) {
  const uint64_t need = noalloc_footprint(m, n);
  if (need > KT_NOALLOC_SCRATCH_BYTES - 4096)
    fail("kt_solve_pairs_device_noalloc: graph too large for the allocation-free solve (m=" + std::to_string(m) +
         ", n=" + std::to_string(n) + " needs up to " + std::to_string(need >> 20) + " MiB of scratch; the limit is " +
         std::to_string(KT_NOALLOC_SCRATCH_BYTES >> 20) + " MiB)");
  if (device < 0 || device >= kMaxDevices) fail("kt_solve_pairs_device_noalloc: device out of range");
  NoAllocDevice& d = g_dev[device];
  std::lock_guard<std::mutex> lock(d.mu);
  set_up(d, device);
  Solver& s = *d.solver;
  KT_HIP(hipSetDevice(device));
  ArenaScope scope(&d.arena);
  s.forget_buffers();
  d.arena.reset();
  s.clear_control();
  s.solve_pairs_device(d_pairs, m, n, res);
  uint64_t count = res->truss_edges;
  if (d_truss) count = s.truss_edges_to_device(d_truss, cap);
  KT_HIP(hipStreamSynchronize(s.stream()));
  return count;
}

}  // namespace kt
