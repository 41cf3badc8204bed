// CPU oracle for the S3TTMc hot path of SymProp.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
// and never by the product library under symprop_amd/.  It restates, in
// plain sequential/OpenMP C++, what the reference computes, from the
// definitions in the reference sources (paths relative to the reference
// checkout):
//
//   combinatorics ......... include/detail/dense/symtensor_base.hpp:12-76
//   packed layout ......... include/detail/dense/symtensor.hpp:58-94,129-142
//   one-sided outer prod .. include/detail/dense/symtensor.hpp:24-39
//   unpack / coeff ........ include/detail/dense/symtensor.hpp:96-126
//   tree contents ......... src/dtree.cpp:38-80 (banerjee_insert)
//   compressed tree ....... src/compressed_dtree.cpp:13-131, hpp:31-50
//   S3TTMc (opt5) ......... src/compressed_dtree.cpp:421-518
//   S3TTMc driver ......... src/sparse/spsymtensor.cpp:30-110
//   norm .................. src/sparse/spsymtensor.cpp:114-138
//   HOEVD Gram ............ src/sparse/spsymtensor.cpp:385-439
//   loader ................ src/utils/loader.cpp:8-36
//   random init ........... src/tucker.cpp:25-36
//
// Pinning (see tests/test_oracle.py): packed layout and outer product by the
// KATs of tests/test_symtensor.cpp:9-58; tree arrays against the reference's
// own CompressedDTree compiled into oracle/_ref (when buildable) and against
// brute-force definitions; S3TTMc against a dense brute force over all
// distinct permutations (an independent restatement of the tensor
// definition), against permanents (tests/test_compressed_dtree.cpp:37-65)
// and against the reference's kronecker_product_opt5 from oracle/_ref.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <numeric>
#include <random>
#include <string>
#include <vector>
#include <omp.h>

namespace {

using u32 = uint32_t;
using u64 = uint64_t;

// symtensor_base.hpp:12-26
u64 binom(u64 n, u64 k) {
  if (k > n) return 0;
  if (k * 2 > n) k = n - k;
  if (k == 0) return 1;
  u64 r = n;
  for (u64 i = 2; i <= k; ++i) {
    r *= (n - i + 1);
    r /= i;
  }
  return r;
}

// symtensor_base.hpp:28-30: number of sorted k-tuples over [0,R)
u64 symsize(u64 R, u64 k) { return binom(k + R - 1, k); }

// symtensor_base.hpp:32-40 (table of 20 entries: 0! .. 19!)
u64 fact(u64 i) {
  u64 f = 1;
  for (u64 t = 2; t <= i; ++t) f *= t;
  return f;
}

// symtensor_base.hpp:59-76 + 51-57: N!/prod(f!) of a sorted tuple
u64 multinomial_sorted(const u32 *t, int n) {
  u64 denom = 1;
  int run = 1;
  for (int j = 1; j <= n; ++j) {
    if (j < n && t[j] == t[j - 1]) {
      ++run;
    } else {
      denom *= fact(run);
      run = 1;
    }
  }
  return fact(n) / denom;
}

// symtensor.hpp:129-142: lexicographic enumeration of sorted k-tuples,
// first index slowest.  Returns the last index of every packed entry.
std::vector<u32> last_index_table(u32 R, int k) {
  std::vector<u32> out;
  out.reserve(symsize(R, k));
  std::vector<u32> idx(k, 0);
  if (k == 0) return out;
  while (true) {
    out.push_back(idx[k - 1]);
    int p = k - 1;
    while (p >= 0 && idx[p] == R - 1) --p;
    if (p < 0) break;
    idx[p]++;
    for (int q = p + 1; q < k; ++q) idx[q] = idx[p];
  }
  return out;
}

// symtensor.hpp:24-39: dst(j1..jk) += src(j1..j_{k-1}) * row(jk), jk >= j_{k-1}
void onesided(double *dst, const double *src, const double *row, u32 R,
              const std::vector<u32> &src_last) {
  u64 res = 0;
  for (u64 s = 0; s < src_last.size(); ++s) {
    const double a = src[s];
    for (u32 l = src_last[s]; l < R; ++l) dst[res++] += a * row[l];
  }
}

bool tuple_less(const u32 *a, const u32 *b, int s) {
  for (int i = 0; i < s; ++i) {
    if (a[i] != b[i]) return a[i] < b[i];
  }
  return false;
}

}  // namespace

struct OrcTree {
  u32 order = 0, dim = 0;
  u64 nnz = 0;
  int depth = 0;                            // order - 1 levels
  std::vector<std::vector<u32>> keys;       // level l: L_l x (l+1) tuples
  std::vector<std::vector<u32>> level_idx;  // level l: last index per node
  std::vector<std::vector<u64>> level_ptr;  // CSR to level l+1 / owners
  std::vector<std::vector<u64>> dep_list;   // level l>=1: L_l x (l+1)
  std::vector<u32> own_rem;
  std::vector<double> own_val;
};

extern "C" {

u64 orc_binom(u64 n, u64 k) { return binom(n, k); }
u64 orc_symtens_size(u64 R, u64 k) { return symsize(R, k); }
u64 orc_multinomial_sorted(const u32 *t, int n) {
  return multinomial_sorted(t, n);
}

// symtensor.hpp:118-126: W[p] = k!/prod(f!) for every packed entry p
void orc_symtens_coeff(u32 R, int k, double *out) {
  std::vector<u32> idx(k, 0);
  u64 p = 0;
  while (true) {
    out[p++] = (double)multinomial_sorted(idx.data(), k);
    int q = k - 1;
    while (q >= 0 && idx[q] == R - 1) --q;
    if (q < 0) break;
    idx[q]++;
    for (int r = q + 1; r < k; ++r) idx[r] = idx[q];
  }
}

// symtensor.hpp:83-94: packed index of a sorted tuple
u64 orc_packed_index(const u32 *t, int k, u32 R) {
  u64 off = 0;
  u32 start = 0;
  for (int j = 0; j < k; ++j) {
    for (u32 i = start; i < t[j]; ++i) off += symsize(R - i, k - 1 - j);
    start = t[j];
  }
  return off;
}

// symtensor.hpp:96-116: packed (P) -> full R^k, every distinct permutation
void orc_unpack(const double *packed, u32 R, int k, double *full) {
  std::vector<u32> idx(k, 0);
  u64 p = 0;
  while (true) {
    std::vector<u32> perm = idx;
    do {
      u64 lin = 0;
      for (int j = 0; j < k; ++j) lin = lin * R + perm[j];
      full[lin] = packed[p];
    } while (std::next_permutation(perm.begin(), perm.end()));
    ++p;
    int q = k - 1;
    while (q >= 0 && idx[q] == R - 1) --q;
    if (q < 0) break;
    idx[q]++;
    for (int r = q + 1; r < k; ++r) idx[r] = idx[q];
  }
}

// symtensor.hpp:24-39 exposed for the KAT of tests/test_symtensor.cpp:26-58
void orc_onesided(double *dst, const double *src, const double *row, u32 R,
                  int src_order) {
  std::vector<u32> last = last_index_table(R, src_order);
  onesided(dst, src, row, R, last);
}

// spsymtensor.cpp:114-138: ||X|| = sqrt(sum v^2 N!/prod f!)
double orc_norm(u32 order, u64 nnz, const u32 *inds, const double *vals) {
  double s = 0.0;
  for (u64 i = 0; i < nnz; ++i) {
    double c = (double)multinomial_sorted(inds + i * order, order);
    s += std::pow(vals[i], 2) * c;
  }
  return std::sqrt(s);
}

// Construction by definition (dtree.cpp:38-80 + compressed_dtree.cpp:13-131):
// level l holds the lexicographically sorted distinct sub-multisets of size
// l+1 of the nonzeros (l+1 <= order-1); level_ptr is the CSR from each node
// to its children (the level-(l+1) nodes with that prefix); the leaf level
// points into the owner list, which for nonzero i and j = N-1..0 holds
// (rem = a_j, value v_i) at leaf M \ {a_j}, ordered by leaf, then nonzero
// input order, then j descending.  dep_list[l][j*(l+1)+k] is the level-(l-1)
// index of node j with position k removed (compressed_dtree.cpp:92-131).
OrcTree *orc_tree_build(u32 order, u32 dim, u64 nnz, const u32 *inds,
                        const double *vals) {
  auto *T = new OrcTree();
  T->order = order;
  T->dim = dim;
  T->nnz = nnz;
  const int N = (int)order;
  const int depth = N - 1;
  T->depth = depth;
  T->keys.resize(depth);
  T->level_idx.resize(depth);
  T->level_ptr.resize(depth);
  T->dep_list.resize(depth);

  // interior levels: all position combinations of size s = l+1
  for (int l = 0; l < depth - 1; ++l) {
    const int s = l + 1;
    std::vector<u32> flat;
    std::vector<int> comb(s);
    for (u64 i = 0; i < nnz; ++i) {
      const u32 *m = inds + i * N;
      for (int t = 0; t < s; ++t) comb[t] = t;
      while (true) {
        for (int t = 0; t < s; ++t) flat.push_back(m[comb[t]]);
        int p = s - 1;
        while (p >= 0 && comb[p] == N - s + p) --p;
        if (p < 0) break;
        comb[p]++;
        for (int q = p + 1; q < s; ++q) comb[q] = comb[q - 1] + 1;
      }
    }
    u64 cnt = flat.size() / s;
    std::vector<u64> ord(cnt);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](u64 a, u64 b) {
      return tuple_less(&flat[a * s], &flat[b * s], s);
    });
    auto &K = T->keys[l];
    for (u64 q = 0; q < cnt; ++q) {
      const u32 *t = &flat[ord[q] * s];
      if (!K.empty() && std::memcmp(&K[K.size() - s], t, s * sizeof(u32)) == 0)
        continue;
      K.insert(K.end(), t, t + s);
    }
  }

  // leaf level and owners
  {
    const int s = depth;
    const u64 no = nnz * (u64)N;
    std::vector<u32> leaf(no * s);
    std::vector<u32> rem(no);
    std::vector<double> val(no);
    for (u64 i = 0; i < nnz; ++i) {
      const u32 *m = inds + i * N;
      for (int j = N - 1, o = 0; j >= 0; --j, ++o) {
        const u64 rec = i * N + o;
        u32 *dst = &leaf[rec * s];
        for (int t = 0, w = 0; t < N; ++t)
          if (t != j) dst[w++] = m[t];
        rem[rec] = m[j];
        val[rec] = vals[i];
      }
    }
    std::vector<u64> ord(no);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](u64 a, u64 b) {
      return tuple_less(&leaf[a * s], &leaf[b * s], s);
    });
    auto &K = T->keys[depth - 1];
    auto &P = T->level_ptr[depth - 1];
    P.push_back(0);
    for (u64 q = 0; q < no; ++q) {
      const u32 *t = &leaf[ord[q] * s];
      if (K.empty() || std::memcmp(&K[K.size() - s], t, s * sizeof(u32)) != 0) {
        if (!K.empty()) P.push_back(q);
        K.insert(K.end(), t, t + s);
      }
      T->own_rem.push_back(rem[ord[q]]);
      T->own_val.push_back(val[ord[q]]);
    }
    P.push_back(no);
  }

  for (int l = 0; l < depth; ++l) {
    const int s = l + 1;
    const auto &K = T->keys[l];
    const u64 L = K.size() / s;
    T->level_idx[l].resize(L);
    for (u64 j = 0; j < L; ++j) T->level_idx[l][j] = K[j * s + s - 1];
  }
  // CSR between interior levels
  for (int l = 0; l < depth - 1; ++l) {
    const int s = l + 1;
    const auto &Kp = T->keys[l];
    const auto &Kc = T->keys[l + 1];
    const u64 Lp = Kp.size() / s, Lc = Kc.size() / (s + 1);
    auto &P = T->level_ptr[l];
    P.assign(Lp + 1, 0);
    u64 parent = 0;
    for (u64 c = 0; c < Lc; ++c) {
      while (std::memcmp(&Kp[parent * s], &Kc[c * (s + 1)], s * sizeof(u32)) != 0)
        ++parent;
      P[parent + 1]++;
    }
    for (u64 p = 0; p < Lp; ++p) P[p + 1] += P[p];
  }
  // exclude-one parents
  for (int l = 1; l < depth; ++l) {
    const int s = l + 1;
    const auto &K = T->keys[l];
    const auto &Kp = T->keys[l - 1];
    const u64 L = K.size() / s, Lp = Kp.size() / (s - 1);
    auto &D = T->dep_list[l];
    D.resize(L * s);
    std::vector<u32> sub(s - 1);
    for (u64 j = 0; j < L; ++j) {
      for (int k = 0; k < s; ++k) {
        for (int t = 0, w = 0; t < s; ++t)
          if (t != k) sub[w++] = K[j * s + t];
        u64 lo = 0, hi = Lp;
        while (lo < hi) {
          u64 mid = (lo + hi) / 2;
          if (tuple_less(&Kp[mid * (s - 1)], sub.data(), s - 1))
            lo = mid + 1;
          else
            hi = mid;
        }
        D[j * s + k] = lo;
      }
    }
  }
  return T;
}

void orc_tree_free(OrcTree *T) { delete T; }
int orc_tree_depth(const OrcTree *T) { return T->depth; }
u64 orc_tree_level_size(const OrcTree *T, int l) {
  return T->level_idx[l].size();
}
u64 orc_tree_num_owners(const OrcTree *T) { return T->own_rem.size(); }

// Copies level l: keys (L x (l+1)), level_idx (L), level_ptr (L+1),
// dep_list (L x (l+1), level >= 1 only).  Null pointers are skipped.
void orc_tree_export(const OrcTree *T, int l, u32 *keys, u32 *level_idx,
                     u64 *level_ptr, u64 *dep_list) {
  if (keys) std::copy(T->keys[l].begin(), T->keys[l].end(), keys);
  if (level_idx)
    std::copy(T->level_idx[l].begin(), T->level_idx[l].end(), level_idx);
  if (level_ptr)
    std::copy(T->level_ptr[l].begin(), T->level_ptr[l].end(), level_ptr);
  if (dep_list && l >= 1)
    std::copy(T->dep_list[l].begin(), T->dep_list[l].end(), dep_list);
}

void orc_tree_owners(const OrcTree *T, u32 *rem, double *val) {
  std::copy(T->own_rem.begin(), T->own_rem.end(), rem);
  std::copy(T->own_val.begin(), T->own_val.end(), val);
}

// compressed_dtree.cpp:421-518 (kronecker_product_opt5) with the caller's
// zeroing of spsymtensor.cpp:91-98: materialised levels, exclude-one
// one-sided outer products, leaf scatter with v*multinomial/N! under a lock
// per Y row, schedule(dynamic,32) over parent nodes.  Y is n x ldy, the first
// P = C(R+N-2, N-1) columns of each row are written.  nthreads <= 0 keeps the
// OpenMP default.
int orc_s3ttmc(const OrcTree *T, const double *U, u32 R, double *Y, u64 ldy,
               int nthreads) {
  const int depth = T->depth;
  const int N = (int)T->order;
  if (depth < 2) return -1;  // order 2: the level loop is empty (A6)
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const u64 n = T->dim;
  std::fill(Y, Y + n * ldy, 0.0);

  std::vector<std::vector<u32>> last(depth + 1);
  for (int k = 1; k <= depth; ++k) last[k] = last_index_table(R, k);

  const u64 P1 = R;
  std::vector<double> parents(T->level_idx[0].size() * P1);
  for (u64 i = 0; i < T->level_idx[0].size(); ++i) {
    const u32 v = T->level_idx[0][i];
    std::copy(U + (u64)(v - 1) * R, U + (u64)v * R, &parents[i * P1]);
  }
  std::vector<omp_lock_t> locks(n);
  for (auto &lk : locks) omp_init_lock(&lk);
  const double nfact = (double)fact(N);

  for (int level = 1; level < depth; ++level) {
    const int s = level + 1;
    const u64 Pc = symsize(R, s);
    const u64 Pp = symsize(R, s - 1);
    const u64 Lc = T->level_idx[level].size();
    const u64 Lp = T->level_idx[level - 1].size();
    const auto &K = T->keys[level];
    const auto &D = T->dep_list[level];
    const auto &CP = T->level_ptr[level - 1];
    const bool leaf = (level == depth - 1);
    std::vector<double> children;
    if (!leaf) children.assign(Lc * Pc, 0.0);

#pragma omp parallel
    {
      std::vector<double> t_row(leaf ? Pc : 0);
      std::vector<u32> full(N);
#pragma omp for schedule(dynamic, 32)
      for (u64 i = 0; i < Lp; ++i) {
        for (u64 j = CP[i]; j < CP[i + 1]; ++j) {
          const u32 *tup = &K[j * s];
          double *dst = leaf ? t_row.data() : &children[j * Pc];
          if (leaf) std::fill(t_row.begin(), t_row.end(), 0.0);
          for (int k = 0; k < s; ++k) {
            const u64 pidx = D[j * s + k];
            onesided(dst, &parents[pidx * Pp], U + (u64)(tup[k] - 1) * R, R,
                     last[s - 1]);
          }
          if (!leaf) continue;
          const auto &OP = T->level_ptr[level];
          for (u64 o = OP[j]; o < OP[j + 1]; ++o) {
            const u32 rem = T->own_rem[o];
            std::copy(tup, tup + s, full.begin());
            full[s] = rem;
            std::sort(full.begin(), full.end());
            const double coeff =
                T->own_val[o] * (double)multinomial_sorted(full.data(), N) / nfact;
            double *yr = Y + (u64)(rem - 1) * ldy;
            omp_set_lock(&locks[rem - 1]);
            for (u64 p = 0; p < Pc; ++p) yr[p] += t_row[p] * coeff;
            omp_unset_lock(&locks[rem - 1]);
          }
        }
      }
    }
    if (!leaf) parents.swap(children);
  }
  for (auto &lk : locks) omp_destroy_lock(&lk);
  return 0;
}

// Dense brute force, independent of the tree: every distinct permutation
// (i1..iN) of every nonzero adds v * prod_b U[i_b, J_b] to the unpacked
// Y[i1, J]; the packed entry p of row i is the unpacked entry at the sorted
// multi-index p.  Cost nnz * N! * R^(N-1); tiny inputs only.
void orc_s3ttmc_dense(u32 order, u32 dim, u64 nnz, const u32 *inds,
                      const double *vals, const double *U, u32 R, double *Y) {
  const int N = (int)order;
  const int k = N - 1;
  u64 Rk = 1;
  for (int t = 0; t < k; ++t) Rk *= R;
  std::vector<double> Yu((u64)dim * Rk, 0.0);
  std::vector<u32> perm(N), J(k);
  for (u64 i = 0; i < nnz; ++i) {
    std::copy(inds + i * N, inds + (i + 1) * N, perm.begin());
    std::sort(perm.begin(), perm.end());
    do {
      for (u64 lin = 0; lin < Rk; ++lin) {
        u64 x = lin;
        double prod = vals[i];
        for (int b = k - 1; b >= 0; --b) {
          J[b] = (u32)(x % R);
          x /= R;
        }
        for (int b = 0; b < k; ++b) prod *= U[(u64)(perm[b + 1] - 1) * R + J[b]];
        Yu[(u64)(perm[0] - 1) * Rk + lin] += prod;
      }
    } while (std::next_permutation(perm.begin(), perm.end()));
  }
  const u64 P = symsize(R, k);
  std::vector<u32> idx(k, 0);
  u64 p = 0;
  while (true) {
    u64 lin = 0;
    for (int b = 0; b < k; ++b) lin = lin * R + idx[b];
    for (u32 r = 0; r < dim; ++r) Y[(u64)r * P + p] = Yu[(u64)r * Rk + lin];
    ++p;
    int q = k - 1;
    while (q >= 0 && idx[q] == R - 1) --q;
    if (q < 0) break;
    idx[q]++;
    for (int t = q + 1; t < k; ++t) idx[t] = idx[q];
  }
}

// HOEVD's A = X_(1) X_(1)^T as spsymtensor.cpp:385-439 builds it: every
// distinct permutation of every nonzero is a COO entry (column = the last N-1
// positions, row = the first), entries are sorted by (column, row), and each
// column adds v1*v2 for every pair k <= l of its entries to A[row1][row2] and,
// off the diagonal, to A[row2][row1].  The column is kept as the index tuple
// itself rather than a base-n integer, so it cannot overflow.  Cost
// nnz * N! entries; tiny inputs only.  A (dim x dim) is overwritten.
void orc_hoevd_gram(u32 order, u32 dim, u64 nnz, const u32 *inds, const double *vals,
                    double *A) {
  const int N = (int)order;
  struct Entry {
    std::vector<u32> col;
    u32 row;
    double v;
  };
  std::vector<Entry> coo;
  std::vector<u32> perm(N);
  for (u64 i = 0; i < nnz; ++i) {
    for (int j = 0; j < N; ++j) perm[j] = inds[i * N + j] - 1;
    std::sort(perm.begin(), perm.end());
    do {
      coo.push_back({std::vector<u32>(perm.begin() + 1, perm.end()), perm[0], vals[i]});
    } while (std::next_permutation(perm.begin(), perm.end()));
  }
  std::sort(coo.begin(), coo.end(), [](const Entry &a, const Entry &b) {
    if (a.col != b.col) return a.col < b.col;
    return a.row < b.row;
  });
  std::fill(A, A + (u64)dim * dim, 0.0);
  u64 i = 0;
  while (i < coo.size()) {
    u64 j = i + 1;
    while (j < coo.size() && coo[j].col == coo[i].col) ++j;
    for (u64 k = i; k < j; ++k)
      for (u64 l = k; l < j; ++l) {
        const u32 r1 = coo[k].row, r2 = coo[l].row;
        const double p = coo[k].v * coo[l].v;
        A[(u64)r1 * dim + r2] += p;
        if (r1 != r2) A[(u64)r2 * dim + r1] += p;
      }
    i = j;
  }
}

// tucker.cpp:30-35: std::mt19937{seed} + normal_distribution<double>(0,1),
// n*R values row-major.  libstdc++'s engine and distribution are called, not
// restated.
void orc_random_factor(u64 seed, u64 count, double *out) {
  std::normal_distribution<double> distr(0.0, 1.0);
  std::mt19937 engine{(std::mt19937::result_type)seed};
  for (u64 i = 0; i < count; ++i) out[i] = distr(engine);
}

// symtens_ttmc.cpp:55-59: std::mt19937(1) + uniform_real_distribution[1,2)
void orc_uniform_factor(u64 seed, u64 count, double *out) {
  std::mt19937 gen((std::mt19937::result_type)seed);
  std::uniform_real_distribution<double> dist(1.0, 2.0);
  for (u64 i = 0; i < count; ++i) out[i] = dist(gen);
}

// loader.cpp:8-36: header "order dim", then N indices and a value per
// line; each tuple is sorted; reading stops at the first failed read.
// Returns nnz, or -1 when the file cannot be opened.  Call with inds/vals
// null to count, then again with buffers.
long long orc_load_tns(const char *path, u32 *order_out, u32 *dim_out,
                       u32 *inds, double *vals) {
  std::ifstream f(path);
  if (!f.is_open()) return -1;
  u32 order = 0, dim = 0;
  f >> order >> dim;
  *order_out = order;
  *dim_out = dim;
  std::vector<u32> t;
  long long cnt = 0;
  while (true) {
    t.clear();
    for (u32 i = 0; i < order; ++i) {
      u32 v;
      f >> v;
      t.push_back(v);
    }
    double v;
    f >> v;
    if (!f.good()) break;
    std::sort(t.begin(), t.end());
    if (inds) std::copy(t.begin(), t.end(), inds + (u64)cnt * order);
    if (vals) vals[cnt] = v;
    ++cnt;
  }
  return cnt;
}

}  // extern "C"
