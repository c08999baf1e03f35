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
#include <atomic>
#include <chrono>
#include <string>
#include <thread>

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

// A thread's 64-byte chunk [c0, c0 + 64) in registers (text is allocated in
// whole tiles, so the load stays inside the buffer; bytes at or past len are
// ignored).
struct Chunk64 {
  uint32_t w[16];
  __device__ __forceinline__ void load(const uint8_t* __restrict__ text, uint64_t c0) {
    const uint4* p = reinterpret_cast<const uint4*>(text + c0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 v = p[k];
      w[4 * k] = v.x;
      w[4 * k + 1] = v.y;
      w[4 * k + 2] = v.z;
      w[4 * k + 3] = v.w;
    }
  }
  __device__ __forceinline__ uint8_t byte(int i) const { return uint8_t(w[i >> 2] >> (8 * (i & 3))); }
};

// Line starts in the chunk: q == 0 or byte q - 1 is '\n', for c0 <= q < c0 + nb
// (nb = bytes of the chunk before len); first: whether c0 itself starts one.
__device__ __forceinline__ uint32_t count_starts(const Chunk64& ch, uint32_t nb, bool first) {
  uint32_t c = first ? 1u : 0u;
#pragma unroll
  for (int i = 0; i < int(kChunkBytes) - 1; ++i) c += (uint32_t(i) + 1 < nb) && ch.byte(i) == '\n';
  return c;
}

__device__ __forceinline__ bool starts_at(const uint8_t* __restrict__ text, uint64_t c0) {
  return c0 == 0 || text[c0 - 1] == '\n';
}

// lines per tile (the workgroup pass of kTileBytes bytes)
__global__ __launch_bounds__(kBlock) void k_tsv_count(const uint8_t* __restrict__ text, uint64_t len,
                                                      uint64_t tiles, uint32_t* tile_cnt) {
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t c0 = t * kTileBytes + uint64_t(threadIdx.x) * kChunkBytes;
    uint32_t c = 0;
    if (c0 < len) {
      Chunk64 ch;
      ch.load(text, c0);
      c = count_starts(ch, uint32_t(min<uint64_t>(kChunkBytes, len - c0)), starts_at(text, c0));
    }
    uint32_t total;
    (void)block_exclusive_scan(c, &total);
    if (threadIdx.x == 0) tile_cnt[t] = total;
  }
}

// Per-line state machine of the contract: 0 before x, 1 in x, 2 between x
// and y, 3 in y, 4 after y. A line's kind at its end: 0 two digit runs
// (states 3, 4), 1 no digit (0), 2 one digit run (1, 2).
struct LineParse {
  uint32_t state = 0, x = 0, y = 0;
  __device__ __forceinline__ void step(uint8_t c) {
    const bool digit = int8_t(c) >= '0';
    const uint32_t d = uint32_t(int8_t(c) - '0');
    switch (state) {
      case 0: if (digit) { state = 1; x = d; } break;
      case 1: if (digit) x = x * 10u + d; else state = 2; break;
      case 2: if (digit) { state = 3; y = d; } break;
      case 3: if (digit) y = y * 10u + d; else state = 4; break;
      default: break;
    }
  }
  __device__ __forceinline__ uint32_t kind() const { return state >= 3 ? 0u : state == 0 ? 1u : 2u; }
};

struct LineOut {
  uint32_t* xs;
  uint32_t* ys;
  uint8_t* kind;
  uint32_t mx = 0, my = 0, blank = 0, single = 0;
  __device__ __forceinline__ void emit(uint64_t at, const LineParse& lp) {
    const uint32_t k = lp.kind();
    xs[at] = k == 0 ? lp.x : 0u;
    ys[at] = k == 0 ? lp.y : 0u;
    kind[at] = uint8_t(k);
    if (k == 0) {
      mx = max(mx, lp.x);
      my = max(my, lp.y);
    }
    blank += k == 1;
    single += k == 2;
  }
};

// Every line starting in a thread's chunk, in file order: xs / ys / kind at
// line index tile_off[t] + (starts before it in the tile). The chunk is
// parsed from registers; the line still open at its end continues from
// memory. stat[0] and stat[1]: largest x and y of two-run lines; stat[2]:
// lines of no digit; stat[3]: lines of one digit run (stat starts at 0).
__global__ __launch_bounds__(kBlock) void k_tsv_parse(const uint8_t* __restrict__ text, uint64_t len, uint64_t tiles,
                                                      const uint64_t* __restrict__ tile_off, uint32_t* xs,
                                                      uint32_t* ys, uint8_t* kind, uint32_t* stat) {
  LineOut out{xs, ys, kind};
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t c0 = t * kTileBytes + uint64_t(threadIdx.x) * kChunkBytes;
    Chunk64 ch;
    uint32_t nb = 0, cnt = 0;
    bool open = false;  // a line that started in this chunk is being parsed
    if (c0 < len) {
      ch.load(text, c0);
      nb = uint32_t(min<uint64_t>(kChunkBytes, len - c0));
      open = starts_at(text, c0);
      cnt = count_starts(ch, nb, open);
    }
    uint32_t total;
    uint64_t at = tile_off[t] + block_exclusive_scan(cnt, &total);
    if (c0 >= len) continue;
    LineParse lp;
#pragma unroll
    for (int i = 0; i < int(kChunkBytes); ++i) {
      if (uint32_t(i) < nb) {
        const uint8_t c = ch.byte(i);
        if (c == '\n') {
          if (open) out.emit(at++, lp);
          lp = LineParse{};
          open = uint32_t(i) + 1 < nb;  // the next line starts here unless the chunk or file ends
        } else if (open) {
          lp.step(c);
        }
      }
    }
    if (open) {  // the last line runs past the chunk: finish it from memory
      uint64_t j = c0 + nb;
      while (j < len && text[j] != '\n') lp.step(text[j++]);
      out.emit(at++, lp);
    }
  }
  // one atomic per wave
  for (int o = kWave / 2; o > 0; o >>= 1) {
    out.mx = max(out.mx, uint32_t(__shfl_xor(out.mx, o)));
    out.my = max(out.my, uint32_t(__shfl_xor(out.my, o)));
    out.blank += uint32_t(__shfl_xor(out.blank, o));
    out.single += uint32_t(__shfl_xor(out.single, o));
  }
  if (lane64() == 0) {
    atomicMax(&stat[0], out.mx);
    atomicMax(&stat[1], out.my);
    if (out.blank) atomicAdd(&stat[2], out.blank);
    if (out.single) atomicAdd(&stat[3], out.single);
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
  // a chunk of the file into a staging buffer, read by a few threads (one
  // core does not keep up with the copy engine)
  auto read_into = [&](uint8_t* dst, uint64_t at, size_t want) {
    constexpr size_t kReaders = 4;
    const size_t part = (want + kReaders - 1) / kReaders;
    std::atomic<bool> ok{true};
    std::thread th[kReaders];
    for (size_t r = 0; r < kReaders; ++r) {
      const size_t a = std::min(want, r * part), e = std::min(want, a + part);
      th[r] = std::thread([&, a, e] {
        for (size_t got = a; got < e;) {
          const ssize_t n = ::pread(file.fd, dst + got, e - got, off_t(at + got));
          if (n <= 0) {
            ok = false;
            return;
          }
          got += size_t(n);
        }
      });
    }
    for (auto& x : th) x.join();
    return ok.load();
  };
  uint64_t off = 0;
  int b = 0;
  bool staged[2] = {false, false};
  while (off < len) {
    if (staged[b]) KT_HIP(hipEventSynchronize(stage_ev_[b]));
    const size_t want = size_t(std::min<uint64_t>(kStageBytes, len - off));
    if (!read_into(stage_[b], off, want)) {
      KT_HIP(hipStreamSynchronize(stream_));
      fail(std::string("cannot read ") + path);
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
