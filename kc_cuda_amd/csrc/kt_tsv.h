#pragma once
#include <atomic>
#include <cstdint>
#include <vector>

namespace kt {

struct ParsedLines {
  std::vector<uint32_t> xs, ys;  // line i: xs[i] goes into row ys[i]
  uint32_t n = 0;                // max id + 1
};

// Throws std::runtime_error on open/map failure or an empty file.
ParsedLines parse_tsv_file(const char* path, unsigned threads = 0);

}  // namespace kt
