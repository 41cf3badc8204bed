// Input side: seeded synthetic generators (the .tns reader is in tns_reader.cpp).
#include "sp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <numeric>
#include <unordered_set>

namespace spamd {

namespace {

u64 fact_u(int n) {
  u64 f = 1;
  for (int i = 2; i <= n; ++i) f *= (u64)i;
  return f;
}

// N!/prod(f!) of a sorted tuple (symtensor_base.hpp:42-76)
u64 multinomial_sorted(const u32 *t, int n) {
  u64 denom = 1;
  int run = 1;
  for (int j = 1; j <= n; ++j) {
    if (j < n && t[j] == t[j - 1]) {
      ++run;
    } else {
      denom *= fact_u(run);
      run = 1;
    }
  }
  return fact_u(n) / denom;
}

struct SplitMix64 {
  u64 x;
  explicit SplitMix64(u64 seed) : x(seed) {}
  u64 next() {
    u64 z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct U128Hash {
  size_t operator()(const u128 &k) const {
    u64 lo = (u64)k, hi = (u64)(k >> 64);
    return (size_t)(lo * 0x9E3779B97F4A7C15ull ^ (hi + 0x632BE59BD9B4E019ull + (lo << 6)));
  }
};

}  // namespace

// ||X|| (src/sparse/spsymtensor.cpp:114-138), summed in input order.
double host_norm(u32 order, u64 nnz, const u32 *inds, const double *vals) {
  double s = 0.0;
  for (u64 i = 0; i < nnz; ++i) {
    const double c = (double)multinomial_sorted(inds + i * order, (int)order);
    s += std::pow(vals[i], 2) * c;
  }
  return std::sqrt(s);
}

}  // namespace spamd

using namespace spamd;

extern "C" {

int sp_generate(int kind, uint32_t order, uint32_t dim, uint64_t nnz, double param,
                uint32_t communities, uint64_t seed, uint32_t *inds, double *vals,
                uint64_t *nnz_out) {
  if (!inds || !vals || !nnz_out || order < 1 || dim < 1 || nnz < 1) return SP_ERR_ARG;
  if (kind < 0 || kind > 2) return SP_ERR_ARG;
  int b = 1;
  while (b < 32 && (1ull << b) <= dim) ++b;
  if ((int)order * b > 128) return SP_ERR_UNSUPPORTED;
  if (kind == 2 && (communities < 1 || communities > dim)) return SP_ERR_ARG;
  try {
    SplitMix64 rng(seed);
    std::vector<double> cdf;
    if (kind == 1) {
      cdf.resize(dim);
      double s = 0.0;
      for (u32 i = 0; i < dim; ++i) {
        s += std::pow((double)(i + 1), -param);
        cdf[i] = s;
      }
      for (auto &x : cdf) x /= s;
    }
    const u32 block = (kind == 2) ? dim / communities : 0;
    std::unordered_set<u128, U128Hash> seen;
    seen.reserve(nnz * 2);
    std::vector<u128> keys;
    std::vector<double> kv;
    keys.reserve(nnz);
    kv.reserve(nnz);
    std::vector<u32> t(order);
    u64 draws = 0;
    const u64 max_draws = nnz * 64 + 1000000;
    while (keys.size() < nnz) {
      if (kind == 0 && draws >= nnz) break;
      if (draws++ > max_draws) return SP_ERR_ARG;  // not enough distinct tuples
      if (kind == 0) {
        for (u32 j = 0; j < order; ++j) t[j] = 1 + (u32)(rng.uniform() * dim);
      } else if (kind == 1) {
        for (u32 j = 0; j < order; ++j) {
          const double u = rng.uniform();
          const u32 k = (u32)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
          t[j] = 1 + std::min(k, dim - 1);
        }
      } else {
        const u32 c = (u32)(rng.uniform() * communities);
        for (u32 j = 0; j < order; ++j) {
          if (rng.uniform() < param)
            t[j] = c * block + 1 + (u32)(rng.uniform() * block);
          else
            t[j] = 1 + (u32)(rng.uniform() * dim);
        }
      }
      for (auto &x : t) x = std::min(std::max(x, 1u), dim);
      std::sort(t.begin(), t.end());
      u128 k = 0;
      for (u32 j = 0; j < order; ++j) k = (k << b) | (u128)t[j];
      const double v = (kind == 1) ? rng.uniform() : 1.0;
      if (!seen.insert(k).second) continue;
      keys.push_back(k);
      kv.push_back(v);
    }
    std::vector<u64> ord(keys.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](u64 a, u64 c) { return keys[a] < keys[c]; });
    const u128 mask = (u128)((1ull << b) - 1);
    for (u64 q = 0; q < ord.size(); ++q) {
      const u128 k = keys[ord[q]];
      for (u32 j = 0; j < order; ++j)
        inds[q * order + j] = (u32)((k >> (b * (order - 1 - j))) & mask);
      vals[q] = kv[ord[q]];
    }
    *nnz_out = keys.size();
    return SP_OK;
  } catch (const std::bad_alloc &) {
    return SP_ERR_OOM;
  }
}

}  // extern "C"
