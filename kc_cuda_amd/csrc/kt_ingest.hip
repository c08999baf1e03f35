// Device-side ingest of the reference's edge-list text: the file's bytes are
// streamed into HBM through pinned staging buffers and parsed there into the
// line arrays the CSR build takes (src/kc_gpu.cpp:243-312 reads and parses
// the file on the host).
//
// Input contract (as the host parser, kt_tsv.cpp): bytes >= '0' (signed
// char) are digits, accumulated as x = 10x + (c - '0') in uint32; line i is
// x<sep>y[rest], x goes into row y, N = max y + 1 (widened past the largest
// x). Lines are cut at '\n'; a line of separators only is skipped. A line
// holding a single digit run takes its y from the next line under the
// contract, which a per-line parse cannot follow: such a file is reported
// back and parsed on the host.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <string>

#include "kt_device.h"
#include "kt_kernels.h"
#include "kt_solver.h"

namespace kt {
namespace {

constexpr uint32_t kChunkBytes = 64;                         // bytes per thread
constexpr uint64_t kTileBytes = uint64_t(kBlock) * kChunkBytes;  // bytes per workgroup pass
constexpr int kIngestGrid = 4096;
constexpr size_t kStageBytes = size_t(64) << 20;             // one pinned staging buffer

__device__ __forceinline__ bool is_digit(uint8_t c) { return int8_t(c) >= '0'; }

struct Fd {
  int fd;
  ~Fd() {
    if (fd >= 0) ::close(fd);
  }
};

// Line starts in [c0, c1): q == 0 or byte q - 1 is '\n' (q < len).
__device__ __forceinline__ uint32_t count_starts(const uint8_t* __restrict__ text, uint64_t len, uint64_t c0) {
  const uint64_t c1 = min(c0 + kChunkBytes, len);
  uint32_t c = 0;
  uint8_t prev = c0 ? text[c0 - 1] : uint8_t('\n');
  for (uint64_t q = c0; q < c1; ++q) {
    c += prev == '\n';
    prev = text[q];
  }
  return c;
}

// lines per tile (the workgroup pass of kTileBytes bytes)
__global__ __launch_bounds__(kBlock) void k_tsv_count(const uint8_t* __restrict__ text, uint64_t len,
                                                      uint64_t tiles, uint32_t* tile_cnt) {
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t c0 = t * kTileBytes + uint64_t(threadIdx.x) * kChunkBytes;
    uint32_t total;
    (void)block_exclusive_scan(c0 < len ? count_starts(text, len, c0) : 0u, &total);
    if (threadIdx.x == 0) tile_cnt[t] = total;
  }
}

// Parses the line starting at q: x, y and its kind (0 two digit runs, 1 no
// digit, 2 one digit run).
__device__ __forceinline__ uint32_t parse_line(const uint8_t* __restrict__ text, uint64_t len, uint64_t q,
                                               uint32_t* x_out, uint32_t* y_out) {
  uint64_t j = q;
  while (j < len && text[j] != '\n' && !is_digit(text[j])) ++j;
  if (j == len || text[j] == '\n') return 1;
  uint32_t x = 0, y = 0;
  while (j < len && is_digit(text[j])) x = x * 10u + uint32_t(int8_t(text[j++]) - '0');
  while (j < len && text[j] != '\n' && !is_digit(text[j])) ++j;
  if (j == len || text[j] == '\n') return 2;
  while (j < len && is_digit(text[j])) y = y * 10u + uint32_t(int8_t(text[j++]) - '0');
  *x_out = x;
  *y_out = y;
  return 0;
}

// Every line starting in a thread's chunk, in file order: xs / ys / kind at
// line index tile_off[t] + (starts before it in the tile). stat[0] and
// stat[1]: largest x and y of two-run lines; stat[2]: lines of no digit;
// stat[3]: lines of one digit run (stat starts at 0).
__global__ __launch_bounds__(kBlock) void k_tsv_parse(const uint8_t* __restrict__ text, uint64_t len, uint64_t tiles,
                                                      const uint64_t* __restrict__ tile_off, uint32_t* xs,
                                                      uint32_t* ys, uint8_t* kind, uint32_t* stat) {
  uint32_t mx = 0, my = 0, blank = 0, single = 0;
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t c0 = t * kTileBytes + uint64_t(threadIdx.x) * kChunkBytes;
    uint32_t total;
    const uint32_t before = block_exclusive_scan(c0 < len ? count_starts(text, len, c0) : 0u, &total);
    if (c0 >= len) continue;
    uint64_t at = tile_off[t] + before;
    const uint64_t c1 = min(c0 + kChunkBytes, len);
    uint8_t prev = c0 ? text[c0 - 1] : uint8_t('\n');
    for (uint64_t q = c0; q < c1; ++q) {
      const bool start = prev == '\n';
      prev = text[q];
      if (!start) continue;
      uint32_t x = 0, y = 0;
      const uint32_t k = parse_line(text, len, q, &x, &y);
      xs[at] = x;
      ys[at] = y;
      kind[at] = uint8_t(k);
      ++at;
      if (k == 0) {
        mx = max(mx, x);
        my = max(my, y);
      }
      blank += k == 1;
      single += k == 2;
    }
  }
  // one atomic per wave
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mx = max(mx, uint32_t(__shfl_xor(mx, o)));
    my = max(my, uint32_t(__shfl_xor(my, o)));
    blank += uint32_t(__shfl_xor(blank, o));
    single += uint32_t(__shfl_xor(single, o));
  }
  if (lane64() == 0) {
    atomicMax(&stat[0], mx);
    atomicMax(&stat[1], my);
    if (blank) atomicAdd(&stat[2], blank);
    if (single) atomicAdd(&stat[3], single);
  }
}

__global__ void k_tsv_flags(const uint8_t* __restrict__ kind, uint64_t L, uint32_t* keep) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < L; i += uint64_t(gridDim.x) * blockDim.x)
    keep[i] = kind[i] == 0;
}

// drops the lines of no digit (keep_off: exclusive scan of the kept flags)
__global__ void k_tsv_compact(const uint32_t* __restrict__ xs, const uint32_t* __restrict__ ys,
                              const uint8_t* __restrict__ kind, const uint64_t* __restrict__ keep_off, uint64_t L,
                              uint32_t* xs2, uint32_t* ys2) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < L; i += uint64_t(gridDim.x) * blockDim.x)
    if (kind[i] == 0) {
      xs2[keep_off[i]] = xs[i];
      ys2[keep_off[i]] = ys[i];
    }
}

}  // namespace

bool Solver::ingest_tsv(const char* path, uint64_t* lines, uint32_t* n) {
  KT_HIP(hipSetDevice(device_));
  const Fd file{::open(path, O_RDONLY)};
  if (file.fd < 0) fail(std::string("cannot open ") + path);
  struct stat st;
  if (::fstat(file.fd, &st) != 0) fail(std::string("cannot stat ") + path);
  const uint64_t len = uint64_t(st.st_size);
  if (len == 0) fail(std::string("empty input file ") + path);
  const uint64_t tiles = (len + kTileBytes - 1) / kTileBytes;
  text_.ensure(tiles * kTileBytes);
  uint8_t* text = text_.as<uint8_t>();
  // file -> pinned staging (two buffers, alternating) -> HBM
  for (int b = 0; b < 2; ++b)
    if (!stage_[b]) KT_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage_[b]), kStageBytes, hipHostMallocDefault));
  uint64_t off = 0;
  int b = 0;
  bool staged[2] = {false, false};
  while (off < len) {
    if (staged[b]) KT_HIP(hipEventSynchronize(stage_ev_[b]));
    const size_t want = size_t(std::min<uint64_t>(kStageBytes, len - off));
    size_t got = 0;
    while (got < want) {
      const ssize_t r = ::pread(file.fd, stage_[b] + got, want - got, off_t(off + got));
      if (r <= 0) {
        KT_HIP(hipStreamSynchronize(stream_));
        fail(std::string("cannot read ") + path);
      }
      got += size_t(r);
    }
    KT_HIP(hipMemcpyAsync(text + off, stage_[b], want, hipMemcpyHostToDevice, stream_));
    KT_HIP(hipEventRecord(stage_ev_[b], stream_));
    staged[b] = true;
    off += want;
    b ^= 1;
  }
  const int grid = int(std::min<uint64_t>(tiles, kIngestGrid));
  tile_cnt_.ensure(sizeof(uint32_t) * (tiles + 1));
  tile_off_.ensure(sizeof(uint64_t) * (tiles + 1));
  k_tsv_count<<<grid, kBlock, 0, stream_>>>(text, len, tiles, tile_cnt_.as<uint32_t>());
  fill_u32(tile_cnt_.as<uint32_t>() + tiles, 1, 0, stream_);
  exclusive_scan_u32_to_u64(tile_cnt_.as<uint32_t>(), tile_off_.as<uint64_t>(), uint32_t(tiles + 1));
  const uint64_t L = read_u64(tile_off_.as<uint64_t>() + tiles);
  lx_.ensure(sizeof(uint32_t) * std::max<uint64_t>(L, 1));
  ly_.ensure(sizeof(uint32_t) * std::max<uint64_t>(L, 1));
  lkind_.ensure(std::max<uint64_t>(L, 1));
  lstat_.ensure(sizeof(uint32_t) * 4);
  fill_u32(lstat_.as<uint32_t>(), 4, 0, stream_);
  k_tsv_parse<<<grid, kBlock, 0, stream_>>>(text, len, tiles, tile_off_.as<uint64_t>(), lx_.as<uint32_t>(),
                                            ly_.as<uint32_t>(), lkind_.as<uint8_t>(), lstat_.as<uint32_t>());
  uint32_t stat[4];
  KT_HIP(hipMemcpyAsync(stat, lstat_.p, sizeof(stat), hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
  const uint32_t mx = stat[0], my = stat[1], blank = stat[2], single = stat[3];
  if (single) return false;  // the contract's cross-line reading: host parser
  uint64_t kept = L;
  if (blank) {
    if (L + 1 >= (uint64_t(1) << 32)) return false;
    kept = L - blank;
    lx2_.ensure(sizeof(uint32_t) * std::max<uint64_t>(kept, 1));
    ly2_.ensure(sizeof(uint32_t) * std::max<uint64_t>(kept, 1));
    lkeep_.ensure(sizeof(uint32_t) * (L + 1));
    lkeep_off_.ensure(sizeof(uint64_t) * (L + 1));
    k_tsv_flags<<<grid_for(L), kBlock, 0, stream_>>>(lkind_.as<uint8_t>(), L, lkeep_.as<uint32_t>());
    fill_u32(lkeep_.as<uint32_t>() + L, 1, 0, stream_);
    exclusive_scan_u32_to_u64(lkeep_.as<uint32_t>(), lkeep_off_.as<uint64_t>(), uint32_t(L + 1));
    k_tsv_compact<<<grid_for(L), kBlock, 0, stream_>>>(lx_.as<uint32_t>(), ly_.as<uint32_t>(), lkind_.as<uint8_t>(),
                                                       lkeep_off_.as<uint64_t>(), L, lx2_.as<uint32_t>(),
                                                       ly2_.as<uint32_t>());
    std::swap(lx_, lx2_);
    std::swap(ly_, ly2_);
  }
  KT_HIP(hipStreamSynchronize(stream_));
  uint64_t nn = kept ? uint64_t(my) + 1 : 0;
  if (kept && mx >= nn) nn = uint64_t(mx) + 1;
  if (nn > 0xFFFFFFFFull) fail("vertex id 2^32 - 1 in the input");
  *lines = kept;
  *n = uint32_t(nn);
  return true;
}

void Solver::copy_ingested(uint32_t* xs, uint32_t* ys, uint64_t count) {
  if (!count) return;
  KT_HIP(hipMemcpyAsync(xs, lx_.p, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipMemcpyAsync(ys, ly_.p, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, stream_));
  KT_HIP(hipStreamSynchronize(stream_));
}

bool Solver::solve_tsv_device(const char* path, kt_result* res, double* parse_ms) {
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t L = 0;
  uint32_t n = 0;
  if (!ingest_tsv(path, &L, &n)) return false;
  if (parse_ms) *parse_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  solve_lines_device(lx_.as<uint32_t>(), ly_.as<uint32_t>(), L, n, res);
  return true;
}

}  // namespace kt
