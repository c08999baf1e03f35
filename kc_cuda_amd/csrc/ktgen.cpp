// Deterministic synthetic graph generators for the k-max-truss workloads.
//
// Every generator returns a simple undirected graph as canonical pairs
// (u < v), sorted by (u, v), without self-loops or duplicates, uint32 ids.
// Randomness comes from a counter-based splitmix64 stream, so the output is a
// pure function of the parameters and the seed, independent of thread count.
//
// The text writer emits the edge list in the input contract of the
// reference's `kc_gpu -f` / `kc_cpu -f` (src/kc_gpu.cpp:243-312): both
// directions of every edge, one `x\ty\n` line each, grouped by the second
// column in ascending order (ties by the first column).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <parallel/algorithm>
#include <vector>
#include <omp.h>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

inline uint64_t fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
// i-th value of the splitmix64 sequence that starts at `key`.
inline uint64_t rnd(uint64_t key, uint64_t i) { return fmix64(key + (i + 1) * kGolden); }
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return fmix64(fmix64(seed * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL) ^ (stream * kGolden));
}
inline double u01(uint64_t r) { return double(r >> 11) * (1.0 / 9007199254740992.0); }

// Sort + unique canonical keys (u << 32 | v, u < v) and unpack to pairs.
// Keys are scattered into 2^16 buckets by their top bits (per-thread
// histograms, so the scatter is stable and deterministic), each bucket is
// sorted and de-duplicated on its own, and the buckets are concatenated.
int64_t finish(std::vector<uint64_t>& keys, uint32_t** out) {
  constexpr int kBits = 16;
  constexpr size_t kB = size_t(1) << kBits;
  const size_t L = keys.size();
  const int nt = omp_get_max_threads();
  std::vector<uint64_t> tmp(L);
  std::vector<size_t> cnt(size_t(nt) * kB, 0), start(kB + 1, 0);
  auto bucket = [](uint64_t k) { return size_t(k >> (64 - kBits)); };
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    const size_t lo = L * t / nt, hi = L * (t + 1) / nt;
    size_t* c = &cnt[size_t(t) * kB];
    for (size_t i = lo; i < hi; ++i) ++c[bucket(keys[i])];
#pragma omp barrier
#pragma omp single
    {
      size_t run = 0;
      for (size_t b = 0; b < kB; ++b) {
        start[b] = run;
        for (int s = 0; s < nt; ++s) {
          const size_t v = cnt[size_t(s) * kB + b];
          cnt[size_t(s) * kB + b] = run;
          run += v;
        }
      }
      start[kB] = run;
    }
    for (size_t i = lo; i < hi; ++i) tmp[c[bucket(keys[i])]++] = keys[i];
  }
  std::vector<size_t> kept(kB + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t b = 0; b < int64_t(kB); ++b) {
    auto first = tmp.begin() + start[b], last = tmp.begin() + start[b + 1];
    std::sort(first, last);
    kept[b] = size_t(std::unique(first, last) - first);
  }
  size_t m = 0;
  for (size_t b = 0; b < kB; ++b) {
    const size_t k = kept[b];
    kept[b] = m;
    m += k;
  }
  // all-ones keys mark dropped samples (self-loops): after de-duplication at
  // most one is left, as the last key of the last bucket
  const size_t last_kept = m - kept[kB - 1];
  if (last_kept && tmp[start[kB - 1] + last_kept - 1] == ~uint64_t(0)) --m;
  std::vector<uint64_t>().swap(keys);
  uint32_t* p = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * 2 * (m > 0 ? m : 1)));
  if (!p) return -1;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t b = 0; b < int64_t(kB); ++b) {
    const size_t o = kept[b], n = (b + 1 < int64_t(kB) ? kept[b + 1] : m) - o;
    for (size_t i = 0; i < n; ++i) {
      const uint64_t k = tmp[start[b] + i];
      p[2 * (o + i)] = uint32_t(k >> 32);
      p[2 * (o + i) + 1] = uint32_t(k);
    }
  }
  *out = p;
  return int64_t(m);
}

inline uint64_t canon(uint32_t a, uint32_t b) {
  return a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
}

// c distinct vertices of [0, n) (Floyd's sampling), in draw order.
std::vector<uint32_t> pick_distinct(uint32_t n, uint32_t c, uint64_t key) {
  std::vector<uint32_t> out;
  if (c > n) c = n;
  std::vector<uint32_t> chosen;
  uint64_t ctr = 0;
  for (uint32_t j = n - c; j < n; ++j) {
    uint32_t t = uint32_t(rnd(key, ctr++) % (uint64_t(j) + 1));
    bool dup = std::find(chosen.begin(), chosen.end(), t) != chosen.end();
    chosen.push_back(dup ? j : t);
  }
  return chosen;
}

void add_clique(std::vector<uint64_t>& keys, const std::vector<uint32_t>& vs) {
  for (size_t i = 0; i < vs.size(); ++i)
    for (size_t j = i + 1; j < vs.size(); ++j) keys.push_back(canon(vs[i], vs[j]));
}

// Uniform random permutation of [0, n) (Fisher-Yates on the stream).
std::vector<uint32_t> permutation(uint32_t n, uint64_t key) {
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = i;
  for (uint32_t i = n; i > 1; --i) {
    uint32_t j = uint32_t(rnd(key, i) % i);
    std::swap(p[i - 1], p[j]);
  }
  return p;
}

// One small graph of a batch, generated serially (the batch runs graphs in
// parallel): ER G(n, avg/(n-1)) or R-MAT on 2^ceil(log2 n) ids with ids >= n
// rejected and a random relabel, plus a planted clique; sorted canonical keys.
std::vector<uint64_t> batch_graph(uint32_t n, bool rmat, double avg_deg, uint32_t clique, uint64_t key) {
  std::vector<uint64_t> keys;
  if (n >= 2 && !rmat) {
    const double p = std::min(1.0, avg_deg / double(n - 1));
    const double lq = std::log1p(-std::min(p, 1.0 - 1e-16));
    uint64_t ctr = 0;
    for (int64_t u = 0; u + 1 < int64_t(n); ++u) {
      int64_t v = u;
      while (p > 0.0) {
        const double r = u01(rnd(key, ctr++));
        v += 1 + ((p >= 1.0) ? 0 : int64_t(std::floor(std::log1p(-r) / lq)));
        if (v >= int64_t(n)) break;
        keys.push_back(canon(uint32_t(u), uint32_t(v)));
      }
    }
  } else if (n >= 2) {
    int scale = 1;
    while ((uint64_t(1) << scale) < n) ++scale;
    const std::vector<uint32_t> perm = permutation(n, fmix64(key ^ 0x5bd1e995ull));
    const uint64_t samples = uint64_t(double(n) * avg_deg / 2.0);
    const uint64_t ta = uint64_t(0.57 * 4294967296.0), tb = uint64_t(0.76 * 4294967296.0),
                   tc = uint64_t(0.95 * 4294967296.0);
    for (uint64_t i = 0; i < samples; ++i) {
      uint32_t u = 0, v = 0;
      uint64_t r = 0;
      for (int lvl = 0; lvl < scale; ++lvl) {
        if ((lvl & 1) == 0) r = rnd(key, i * 32 + uint64_t(lvl >> 1));
        const uint64_t x = (lvl & 1) ? (r >> 32) : (r & 0xffffffffULL);
        const uint32_t bit = uint32_t(1) << (scale - 1 - lvl);
        if (x < ta) {
        } else if (x < tb) {
          v |= bit;
        } else if (x < tc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      if (u >= n || v >= n || u == v) continue;
      keys.push_back(canon(perm[u], perm[v]));
    }
  }
  const std::vector<uint32_t> cl = pick_distinct(n, std::min(clique, n), fmix64(key ^ 0x9e3779b9ull));
  add_clique(keys, cl);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  return keys;
}

}  // namespace

extern "C" {

void ktgen_free(void* p) { std::free(p); }

// A batch of `count` independent graphs: n log-uniform in [n_min, n_max], ER or
// R-MAT (a, b, c) = (.57, .19, .19) with equal odds, average degree uniform in
// [deg_min, deg_max], planted clique uniform in [clique_min, clique_max]
// (capped at n). Outputs, malloc'd: pairs (canonical, sorted within each
// graph, graphs concatenated), offsets (count + 1 pair offsets) and n per
// graph. Returns the total number of pairs, or -1.
int64_t ktgen_batch(uint32_t count, uint32_t n_min, uint32_t n_max, double deg_min, double deg_max,
                    uint32_t clique_min, uint32_t clique_max, uint64_t seed, uint32_t** pairs_out,
                    uint64_t** off_out, uint32_t** n_out) {
  std::vector<std::vector<uint64_t>> gs(count);
  std::vector<uint32_t> ns(count);
  const double l0 = std::log(double(n_min)), l1 = std::log(double(n_max));
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < int64_t(count); ++i) {
    const uint64_t key = stream_key(seed, 100 + uint64_t(i));
    const double a = u01(rnd(key, 0)), b = u01(rnd(key, 1)), c = u01(rnd(key, 2)), d = u01(rnd(key, 3));
    uint32_t n = uint32_t(std::llround(std::exp(l0 + (l1 - l0) * a)));
    n = std::min(std::max(n, n_min), n_max);
    const bool rmat = b < 0.5;
    const double deg = deg_min + (deg_max - deg_min) * c;
    const uint32_t cl = clique_min + uint32_t(d * double(clique_max - clique_min + 1));
    ns[size_t(i)] = n;
    gs[size_t(i)] = batch_graph(n, rmat, deg, std::min(cl, clique_max), fmix64(key ^ 0xabcdefull));
  }
  uint64_t total = 0;
  uint64_t* off = static_cast<uint64_t*>(std::malloc(sizeof(uint64_t) * (size_t(count) + 1)));
  uint32_t* nn = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (count ? count : 1)));
  if (!off || !nn) return -1;
  for (uint32_t i = 0; i < count; ++i) {
    off[i] = total;
    nn[i] = ns[i];
    total += gs[i].size();
  }
  off[count] = total;
  uint32_t* p = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * 2 * (total ? total : 1)));
  if (!p) return -1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < int64_t(count); ++i) {
    const auto& k = gs[size_t(i)];
    for (size_t j = 0; j < k.size(); ++j) {
      p[2 * (off[i] + j)] = uint32_t(k[j] >> 32);
      p[2 * (off[i] + j) + 1] = uint32_t(k[j]);
    }
  }
  *pairs_out = p;
  *off_out = off;
  *n_out = nn;
  return int64_t(total);
}

// Erdos-Renyi G(n, p) with p = avg_deg / (n - 1), by geometric skipping per
// row, plus a planted clique on `clique` distinct random vertices.
int64_t ktgen_er_clique(uint32_t n, double avg_deg, uint32_t clique, uint64_t seed, uint32_t** out) {
  if (n < 2) {
    std::vector<uint64_t> none;
    return finish(none, out);
  }
  const double p = std::min(1.0, avg_deg / double(n - 1));
  const uint64_t k_rows = stream_key(seed, 1);
  const int nt = omp_get_max_threads();
  std::vector<std::vector<uint64_t>> parts(nt);
#pragma omp parallel
  {
    auto& mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
    for (int64_t u = 0; u < int64_t(n) - 1; ++u) {
      const uint64_t key = stream_key(k_rows, uint64_t(u));
      uint64_t ctr = 0;
      int64_t v = u;  // next candidate is v + 1 + skip
      if (p <= 0.0) continue;
      const double lq = std::log1p(-std::min(p, 1.0 - 1e-16));
      while (true) {
        double r = u01(rnd(key, ctr++));
        int64_t skip = (p >= 1.0) ? 0 : int64_t(std::floor(std::log1p(-r) / lq));
        v += 1 + skip;
        if (v >= int64_t(n)) break;
        mine.push_back(canon(uint32_t(u), uint32_t(v)));
      }
    }
  }
  std::vector<uint64_t> keys;
  size_t tot = 0;
  for (auto& q : parts) tot += q.size();
  keys.reserve(tot + size_t(clique) * clique / 2);
  for (auto& q : parts) keys.insert(keys.end(), q.begin(), q.end());
  add_clique(keys, pick_distinct(n, clique, stream_key(seed, 2)));
  return finish(keys, out);
}

// R-MAT (Chakrabarti et al.): 2^scale vertices, edge_factor * 2^scale
// samples with quadrant probabilities (a, b, c, 1-a-b-c); self-loops dropped,
// symmetrised, deduplicated, then (if relabel) a uniform random vertex
// permutation; plus an optional planted clique on random vertices.
int64_t ktgen_rmat(int scale, int edge_factor, double a, double b, double c, int relabel,
                   uint32_t clique, uint64_t seed, uint32_t** out) {
  const uint32_t n = uint32_t(1) << scale;
  const int64_t samples = int64_t(edge_factor) << scale;
  const uint64_t ta = uint64_t(a * 4294967296.0);
  const uint64_t tb = uint64_t((a + b) * 4294967296.0);
  const uint64_t tc = uint64_t((a + b + c) * 4294967296.0);
  const uint64_t key = stream_key(seed, 1);
  std::vector<uint32_t> perm;
  if (relabel) perm = permutation(n, stream_key(seed, 3));
  std::vector<uint64_t> keys(static_cast<size_t>(samples));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < samples; ++i) {
    uint32_t u = 0, v = 0;
    uint64_t r = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 1) == 0) r = rnd(key, uint64_t(i) * 32 + uint64_t(lvl >> 1));
      const uint64_t x = (lvl & 1) ? (r >> 32) : (r & 0xffffffffULL);
      const uint32_t bit = uint32_t(1) << (scale - 1 - lvl);
      if (x < ta) {
      } else if (x < tb) {
        v |= bit;
      } else if (x < tc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
    if (relabel) {
      u = perm[u];
      v = perm[v];
    }
    keys[size_t(i)] = (u == v) ? ~uint64_t(0) : canon(u, v);
  }
  if (clique) add_clique(keys, pick_distinct(n, clique, stream_key(seed, 2)));
  return finish(keys, out);
}

// Chung-Lu power-law graph: weights w_i = (i+1)^(-1/(gamma-1)), scaled so the
// expected degree sum is 2 * samples, capped at sqrt(sum w); `samples`
// endpoint pairs drawn independently proportional to weight (Walker/Vose
// alias table), optionally relabelled by a random permutation (hubs are not
// left at the small ids), self-loops dropped, deduplicated, plus a planted
// clique.
int64_t ktgen_chung_lu(uint32_t n, double gamma, int64_t samples, int relabel, uint32_t clique, uint64_t seed,
                       uint32_t** out) {
  if (n < 2 || samples <= 0) {
    std::vector<uint64_t> none;
    return finish(none, out);
  }
  std::vector<double> w(n);
  const double ex = -1.0 / (gamma - 1.0);
  double sum = 0;
  for (uint32_t i = 0; i < n; ++i) {
    w[i] = std::pow(double(i) + 1.0, ex);
    sum += w[i];
  }
  const double scale = 2.0 * double(samples) / sum;
  double tot = 0;
  for (uint32_t i = 0; i < n; ++i) {
    w[i] *= scale;
    tot += w[i];
  }
  const double cap = std::sqrt(tot);
  tot = 0;
  for (uint32_t i = 0; i < n; ++i) {
    w[i] = std::min(w[i], cap);
    tot += w[i];
  }
  // Vose alias table
  std::vector<float> prob(n);
  std::vector<uint32_t> alias(n);
  {
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
      q[i] = w[i] * double(n) / tot;
      (q[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t s = small.back(), l = large.back();
      small.pop_back();
      prob[s] = float(q[s]);
      alias[s] = l;
      q[l] = (q[l] + q[s]) - 1.0;
      if (q[l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (uint32_t i : large) prob[i] = 1.0f, alias[i] = i;
    for (uint32_t i : small) prob[i] = 1.0f, alias[i] = i;
  }
  std::vector<double>().swap(w);
  const uint64_t key = stream_key(seed, 1);
  std::vector<uint32_t> perm;
  if (relabel) perm = permutation(n, stream_key(seed, 3));
  std::vector<uint64_t> keys(static_cast<size_t>(samples));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < samples; ++i) {
    uint32_t e[2];
    for (int s = 0; s < 2; ++s) {
      uint64_t r = rnd(key, uint64_t(i) * 2 + s);
      uint32_t col = uint32_t((r >> 32) * uint64_t(n) >> 32);
      float f = float(uint32_t(r) >> 8) * (1.0f / 16777216.0f);
      e[s] = (f < prob[col]) ? col : alias[col];
      if (relabel) e[s] = perm[e[s]];
    }
    keys[size_t(i)] = (e[0] == e[1]) ? ~uint64_t(0) : canon(e[0], e[1]);
  }
  if (clique) add_clique(keys, pick_distinct(n, clique, stream_key(seed, 2)));
  return finish(keys, out);
}

// Writes the reference's TSV input contract: both directions, grouped by the
// second column ascending, ties by the first column. Returns bytes written or
// -1 on I/O failure.
int64_t ktgen_write_tsv(const char* path, const uint32_t* pairs, int64_t m) {
  std::vector<uint64_t> lines(size_t(2 * (m > 0 ? m : 0)));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    const uint32_t u = pairs[2 * i], v = pairs[2 * i + 1];
    lines[size_t(2 * i)] = uint64_t(v) << 32 | u;      // line "u v": row v
    lines[size_t(2 * i + 1)] = uint64_t(u) << 32 | v;  // line "v u": row u
  }
  __gnu_parallel::sort(lines.begin(), lines.end());
  FILE* f = std::fopen(path, "wb");
  if (!f) return -1;
  // format blocks of lines in parallel, write them in order
  const size_t L = lines.size(), block = 1 << 20;
  const size_t nb = (L + block - 1) / block;
  int64_t total = 0;
  const size_t batch = 64;
  std::vector<std::vector<char>> bufs(batch);
  for (size_t b0 = 0; b0 < nb; b0 += batch) {
    const size_t b1 = std::min(nb, b0 + batch);
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t b = b0; b < b1; ++b) {
      auto& buf = bufs[b - b0];
      buf.resize(block * 24);
      size_t pos = 0;
      char tmp[12];
      for (size_t i = b * block; i < std::min(L, (b + 1) * block); ++i) {
        for (int half = 0; half < 2; ++half) {
          uint32_t x = half ? uint32_t(lines[i] >> 32) : uint32_t(lines[i]);
          int t = 0;
          do { tmp[t++] = char('0' + x % 10); x /= 10; } while (x);
          while (t) buf[pos++] = tmp[--t];
          buf[pos++] = half ? '\n' : '\t';
        }
      }
      buf.resize(pos);
    }
    for (size_t b = b0; b < b1; ++b) {
      auto& buf = bufs[b - b0];
      if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
        std::fclose(f);
        return -1;
      }
      total += int64_t(buf.size());
    }
  }
  if (std::fclose(f) != 0) return -1;
  return total;
}

}  // extern "C"
