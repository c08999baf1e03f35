// Parallel host parser for the reference's edge-list text.
//
// Input contract (src/kc_gpu.cpp:243-312): each line is x<sep>y[rest];
// separators are bytes < '0'; digits are bytes >= '0' accumulated as
// x = 10x + (c - '0'); anything after y up to the newline is ignored. Line x y
// puts x into row y; N = max(y) + 1. Chunks are split on newline boundaries
// and parsed by std::thread workers into private buffers, then concatenated
// in file order, so the result does not depend on the thread count (the
// reference's fixed per-thread regions do, see its OMP_NUM_THREADS limits).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "kt_tsv.h"

namespace kt {

namespace {

struct Chunk {
  std::vector<uint32_t> xs, ys;
  uint32_t maxy = 0;
  bool any = false;
};

void parse_range(const char* d, uint64_t len, uint64_t j, uint64_t r, Chunk* out) {
  // j and r are line starts (or len); lines are [j, r)
  out->xs.reserve((r - j) / 8 + 16);
  out->ys.reserve((r - j) / 8 + 16);
  while (j < r && !(d[j] >= '0')) ++j;
  while (j < r) {
    uint32_t x = 0, y = 0;
    while (j < len && d[j] >= '0') x = x * 10u + uint32_t(d[j++] - '0');
    while (j < len && !(d[j] >= '0')) ++j;
    while (j < len && d[j] >= '0') y = y * 10u + uint32_t(d[j++] - '0');
    while (j < r && d[j] != '\n') ++j;
    while (j < r && !(d[j] >= '0')) ++j;
    out->xs.push_back(x);
    out->ys.push_back(y);
    if (!out->any || y > out->maxy) out->maxy = y;
    out->any = true;
  }
}

}  // namespace

ParsedLines parse_tsv_file(const char* path, unsigned threads) {
  int fd = ::open(path, O_RDONLY);
  if (fd < 0) throw std::runtime_error(std::string("cannot open ") + path);
  struct stat st;
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    throw std::runtime_error(std::string("cannot stat ") + path);
  }
  const uint64_t len = uint64_t(st.st_size);
  ParsedLines out;
  if (len == 0) {
    ::close(fd);
    throw std::runtime_error(std::string("empty input file ") + path);
  }
  void* map = ::mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (map == MAP_FAILED) throw std::runtime_error(std::string("cannot map ") + path);
  const char* d = static_cast<const char*>(map);
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  const uint64_t want = std::max<uint64_t>(1, std::min<uint64_t>(threads * 4ull, len / (1 << 16) + 1));
  // chunk boundaries at line starts
  std::vector<uint64_t> cut(want + 1);
  cut[0] = 0;
  for (uint64_t c = 1; c < want; ++c) {
    uint64_t p = std::max(cut[c - 1], len * c / want);
    while (p < len && d[p - 1] != '\n') ++p;
    cut[c] = p;
  }
  cut[want] = len;
  std::vector<Chunk> chunks(want);
  std::vector<std::thread> pool;
  std::atomic<uint64_t> next{0};
  for (unsigned t = 0; t < threads; ++t) {
    pool.emplace_back([&]() {
      for (uint64_t c = next++; c < want; c = next++)
        if (cut[c] < cut[c + 1]) parse_range(d, len, cut[c], cut[c + 1], &chunks[c]);
    });
  }
  for (auto& th : pool) th.join();
  ::munmap(map, len);
  uint64_t total = 0;
  uint32_t maxy = 0;
  bool any = false;
  std::vector<uint64_t> at(want + 1, 0);
  for (uint64_t c = 0; c < want; ++c) {
    at[c] = total;
    total += chunks[c].xs.size();
    if (chunks[c].any) {
      maxy = any ? std::max(maxy, chunks[c].maxy) : chunks[c].maxy;
      any = true;
    }
  }
  out.xs.resize(total);
  out.ys.resize(total);
  pool.clear();
  next = 0;
  for (unsigned t = 0; t < threads; ++t) {
    pool.emplace_back([&]() {
      for (uint64_t c = next++; c < want; c = next++) {
        std::copy(chunks[c].xs.begin(), chunks[c].xs.end(), out.xs.begin() + at[c]);
        std::copy(chunks[c].ys.begin(), chunks[c].ys.end(), out.ys.begin() + at[c]);
        std::vector<uint32_t>().swap(chunks[c].xs);
        std::vector<uint32_t>().swap(chunks[c].ys);
      }
    });
  }
  for (auto& th : pool) th.join();
  out.n = any ? maxy + 1 : 0;
  // ids beyond the largest row index have no row; treat them like the
  // reference does not (it would index past deg_origin): widen n instead.
  uint32_t mx = 0;
  for (uint32_t x : out.xs) mx = std::max(mx, x);
  if (!out.xs.empty() && mx >= out.n) out.n = mx + 1;
  return out;
}

}  // namespace kt
