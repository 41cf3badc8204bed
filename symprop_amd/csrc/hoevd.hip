// HOEVD initial factor: top-R eigenvectors of A = X_(1) X_(1)^T
// (src/sparse/spsymtensor.cpp:363-382, sparse_syrk 385-409, convert_to_coo
// 411-439).
//
// The reference materialises X_(1) as a COO over every distinct permutation of
// every nonzero (column = base-n key of the last N-1 positions) and
// accumulates the products of entries sharing a column.  A column is an
// ordering of an (N-1)-sub-multiset S of a nonzero, i.e. a leaf of the
// compressed dimension tree, and the entries of that column are the leaf's
// owners (rem, v) with repeated owners (same nonzero, repeated value) counted
// once.  Every ordering of S carries the same entries, so
//
//   A = sum_S  mult(S) * s_S s_S^T,   mult(S) = (N-1)! / prod f_S!,
//
// with s_S the sparse vector of the leaf's distinct owners.  No permutation
// list, no key overflow for n^(N-1) >= 2^64.
//
// Two solvers, chosen by dim (hoevd_init):
//  - dense (dim <= 4096): row r of A is produced by one wavefront from the
//    owners whose rem is r (the rem-sorted work list, leaves ascending),
//    scattering into an LDS tile of A[r, c0 : c0+TILE].  Contributions are
//    added in ascending leaf order, each as mult * (v_r * v_c), so A[r][c]
//    and A[c][r] are the same floating-point sum and A is exactly symmetric;
//    then the top-R eigen-step on A.
//  - block Krylov (larger dim): A is never formed; HoevdOperator applies it
//    to blocks of vectors from the same owner lists and hoevd_krylov runs
//    block Lanczos with full reorthogonalisation on it.
// Both are deterministic.
#include "sp_internal.hpp"
#include "wave.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace spamd {

rocblas_handle blas_handle(Ctx &c);

namespace {

constexpr int GRAM_TILE = 4096;  // columns of A per workgroup (32 KB of LDS)
constexpr int PRE = 4;           // owners per leaf staged in registers

__device__ inline double leaf_multiplicity(const u32 *t, int m) {
  // m! / prod f! of the sorted leaf tuple
  double fact[20];
  fact[0] = 1.0;
  for (int i = 1; i < 20; ++i) fact[i] = fact[i - 1] * (double)i;
  double d = 1.0;
  int run = 1;
  for (int i = 1; i <= m; ++i) {
    if (i < m && t[i] == t[i - 1]) {
      ++run;
    } else {
      d *= fact[run];
      run = 1;
    }
  }
  return fact[m] / d;
}

// grid: (dim rows, ceil(dim / GRAM_TILE) column tiles); block: one wavefront.
__global__ __launch_bounds__(64) void hoevd_gram_rows(
    const u64 *__restrict__ segptr, const u32 *__restrict__ w_leaf,
    const double *__restrict__ w_val, const u32 *__restrict__ w_tuple, int m,
    const u64 *__restrict__ leaf_ptr, const u32 *__restrict__ own_rem,
    const double *__restrict__ own_val, u32 n, double *__restrict__ A) {
  __shared__ double acc[GRAM_TILE];
  const int lane = threadIdx.x;
  const u32 r = blockIdx.x;
  const u32 c0 = blockIdx.y * GRAM_TILE;
  const u32 cw = min((u32)GRAM_TILE, n - c0);
  for (u32 i = lane; i < cw; i += 64) acc[i] = 0.0;
  __syncthreads();
  const u64 q0 = segptr[r], q1 = segptr[r + 1];
  for (u64 qb = q0; qb < q1; qb += 64) {
    // stage: lane l reads owner qb + l of row r and up to PRE owners of its leaf
    const u64 q = qb + lane;
    const bool valid = q < q1;
    u32 leaf = 0, k = 0;
    u64 p0 = 0;
    double vr = 0.0, mult = 0.0;
    u32 prem[PRE];
    double pval[PRE];
    if (valid) {
      leaf = w_leaf[q];
      const bool dup = q > q0 && w_leaf[q - 1] == leaf;
      vr = w_val[q];
      p0 = leaf_ptr[leaf];
      k = dup ? 0u : (u32)(leaf_ptr[leaf + 1] - p0);
      u32 t[16];
      for (int i = 0; i < m; ++i) t[i] = w_tuple[q * m + i];
      mult = leaf_multiplicity(t, m);
    }
#pragma unroll
    for (int i = 0; i < PRE; ++i) {
      const bool in = (u32)i < k;
      prem[i] = in ? own_rem[p0 + i] : 0u;
      pval[i] = in ? own_val[p0 + i] : 0.0;
    }
    const int nb = (int)min((u64)64, q1 - qb);
    for (int j = 0; j < nb; ++j) {
      const u32 kj = __shfl(k, j);
      if (kj == 0) continue;  // duplicate owner of the row (uniform branch)
      const double wj = __shfl(mult, j);
      const double vj = __shfl(vr, j);
      if (kj <= (u32)PRE) {
        // lane i < kj takes staged owner i of lane j's leaf
        u32 rem = 0, prev = 0;
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < PRE; ++i) {
          const u32 ri = __shfl(prem[i], j);
          const double vi = __shfl(pval[i], j);
          if (lane == i) { rem = ri; v = vi; }
          if (lane == i + 1) prev = ri;
        }
        if ((u32)lane < kj && !(lane > 0 && prev == rem)) {
          const u32 col = rem - 1;
          if (col >= c0 && col < c0 + cw) acc[col - c0] += wj * (vj * v);
        }
      } else {
        const u64 pj = __shfl(p0, j);
        for (u64 p = pj + lane; p < pj + kj; p += 64) {
          const u32 rem = own_rem[p];
          if (p > pj && own_rem[p - 1] == rem) continue;  // repeated owner
          const u32 col = rem - 1;
          if (col >= c0 && col < c0 + cw) acc[col - c0] += wj * (vj * own_val[p]);
        }
      }
      __syncthreads();  // the next leaf may update the same columns
    }
  }
  __syncthreads();
  double *row = A + (u64)r * n + c0;
  for (u32 i = lane; i < cw; i += 64) row[i] = acc[i];
}

__global__ void transpose_cols(const double *__restrict__ Z, u64 n, u32 R,
                               double *__restrict__ U) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * R) return;
  const u64 i = t / R;
  const u32 j = (u32)(t % R);
  U[t] = Z[(u64)j * n + i];
}

// ---- A V without forming A (the top-R solver for large dim) ----------------
//
// A leaf S with a single distinct owner (rem r, value v) adds mult(S) v^2 to
// A[r][r] and nothing else, so A = D + sum over the other leaves, with the
// diagonal D computed once.  For the rest, A V is taken 16 columns of V at a
// time:
//   stage 1  C[S, :] = mult(S) * sum over the distinct owners (rem, v) of leaf S
//            of v * V[rem-1, :]           (owners in leaf order, level_ptr)
//   stage 2  Y[r, :] = D[r] V[r, :] + sum over the distinct owners (leaf S, v)
//            of row r of v * C[S, :]      (owners in rem order: the work list
//                                          without the single-owner leaves)
// Both stages are segmented sums of gathered 16-wide rows.  Leaves with at
// most kLightOwners owners take 16 lanes each; longer segments are cut into
// pieces of at most kPieceOwners owners, one wavefront per piece (four
// 16-lane groups over every fourth owner), whose partial rows seg_finish adds
// in piece order.  Every sum has a fixed order: the result is deterministic.
constexpr int kCols = 16;
constexpr u64 kLightOwners = 32;
constexpr u64 kPieceOwners = 256;

template <class K>
__global__ void leaf_mult(const K *__restrict__ keys, u64 L, int m, int b,
                          double *__restrict__ mult) {
  const u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L) return;
  const K key = keys[s];
  const K mask = (K)((1ull << b) - 1);
  u32 t[16];
  for (int i = 0; i < m; ++i) t[i] = (u32)((key >> (b * (m - 1 - i))) & mask);
  mult[s] = leaf_multiplicity(t, m);
}

// Leaf kinds: 0 one distinct owner (diagonal only), 1 several, at most
// kLightOwners owner entries, 2 more entries (cut into pieces).  light and
// heavy get 0/1 flags for the compaction scans, with a 0 after the last leaf.
__global__ void leaf_kinds(const u64 *__restrict__ ptr, u64 L, const u32 *__restrict__ own_rem,
                           unsigned char *__restrict__ kind, u64 *__restrict__ light,
                           u64 *__restrict__ heavy) {
  const u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > L) return;
  if (s == L) {
    light[s] = heavy[s] = 0;
    return;
  }
  const u64 p0 = ptr[s], p1 = ptr[s + 1];
  int k = 2;
  if (p1 - p0 <= kLightOwners) {
    k = 0;
    for (u64 p = p0 + 1; p < p1; ++p)
      if (own_rem[p] != own_rem[p - 1]) k = 1;
  }
  kind[s] = (unsigned char)k;
  light[s] = k == 1;
  heavy[s] = k == 2;
}

// list[pos[s]] = s for the flagged s < L
__global__ void compact_ids(const u64 *__restrict__ flag, const u64 *__restrict__ pos, u64 L,
                            u32 *__restrict__ list) {
  const u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < L && flag[s]) list[pos[s]] = (u32)s;
}

// the owner range of each listed leaf
__global__ void leaf_ranges(const u32 *__restrict__ list, u64 nl, const u64 *__restrict__ ptr,
                            u64 *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  out[2 * t] = ptr[list[t]];
  out[2 * t + 1] = ptr[list[t] + 1];
}

// work-list entries whose leaf has several owners (flag, with a 0 after the last)
__global__ void entry_flags(const u32 *__restrict__ w_leaf, u64 E,
                            const unsigned char *__restrict__ kind, u64 *__restrict__ flag) {
  const u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < E) flag[q] = kind[w_leaf[q]] != 0;
  else if (q == E) flag[q] = 0;
}

__global__ void compact_entries(const u64 *__restrict__ flag, const u64 *__restrict__ pos, u64 E,
                                const u32 *__restrict__ w_leaf, const double *__restrict__ w_val,
                                u32 *__restrict__ leaf2, double *__restrict__ val2) {
  const u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < E && flag[q]) {
    leaf2[pos[q]] = w_leaf[q];
    val2[pos[q]] = w_val[q];
  }
}

// seg2[r] = pos[segptr[r]], r <= n: the row pointer of the compacted work list
__global__ void remap_rows(const u64 *__restrict__ segptr, const u64 *__restrict__ pos, u64 n,
                           u64 *__restrict__ seg2) {
  const u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r <= n) seg2[r] = pos[segptr[r]];
}

// stage 1 for the light leaves with several owners: 16 lanes per listed leaf,
// C row-major L x 16 (rows of the other leaves are not touched)
__global__ void leaf_rows_light(const u32 *__restrict__ list, u64 nl, const u64 *__restrict__ ptr,
                                const u32 *__restrict__ own_rem,
                                const double *__restrict__ own_val,
                                const double *__restrict__ mult, const double *__restrict__ V,
                                int ldv, int col0, double *__restrict__ C) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 i = t / kCols;
  const int j = (int)(t % kCols);
  if (i >= nl) return;
  const u64 s = list[i];
  const u64 p0 = ptr[s], p1 = ptr[s + 1];
  double acc = 0.0;
  for (u64 p = p0; p < p1; ++p) {
    const u32 rem = own_rem[p];
    if (p > p0 && own_rem[p - 1] == rem) continue;  // repeated owner
    acc += own_val[p] * V[(u64)(rem - 1) * ldv + col0 + j];
  }
  C[s * kCols + j] = mult[s] * acc;
}

struct SegPiece {
  u64 o0, o1;  // owner range of the piece
  u64 first;   // first owner of its segment (the repeated-owner test)
};

// one wavefront per piece: part[piece, :] = sum over its distinct owners q of
// val[q] * X[idx[q] - base, xcol0 : xcol0 + 16]
__global__ void __launch_bounds__(64)
seg_pieces(const SegPiece *__restrict__ pieces, const u32 *__restrict__ idx, u32 base,
           const double *__restrict__ val, const double *__restrict__ X, int ldx, int xcol0,
           double *__restrict__ part) {
  const SegPiece pc = pieces[blockIdx.x];
  const int lane = threadIdx.x;
  const int g = lane / kCols, j = lane % kCols;
  double acc = 0.0;
  for (u64 q = pc.o0 + g; q < pc.o1; q += 64 / kCols) {
    const u32 id = idx[q];
    if (q > pc.first && idx[q - 1] == id) continue;  // repeated owner
    acc += val[q] * X[(u64)(id - base) * ldx + xcol0 + j];
  }
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (g == 0) part[(u64)blockIdx.x * kCols + j] = acc;
}

struct SegOut {
  u64 row;     // output row
  u64 p0, p1;  // its pieces
};

// out[row, ocol0 + j] = (scale ? scale[row] : 1) * (sum of its pieces, in
// order) [+ diag[row] * X[row, ocol0 + j] when diag is given; X: ld ldo]
__global__ void seg_finish(const SegOut *__restrict__ segs, u64 nseg,
                           const double *__restrict__ part, const double *__restrict__ scale,
                           const double *__restrict__ diag, const double *__restrict__ X,
                           double *__restrict__ out, int ldo, int ocol0) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 s = t / kCols;
  const int j = (int)(t % kCols);
  if (s >= nseg) return;
  const SegOut so = segs[s];
  double acc = 0.0;
  for (u64 p = so.p0; p < so.p1; ++p) acc += part[p * kCols + j];
  if (scale) acc *= scale[so.row];
  const u64 o = so.row * ldo + ocol0 + j;
  if (diag) acc += diag[so.row] * X[o];
  out[o] = acc;
}

// D[r] = sum over row r's distinct single-owner leaves of mult * v^2.  The
// rem-sorted work list is cut into pieces of at most kDiagPieceOwners
// entries (never across rows); one wavefront sums a piece (lanes over its
// entries, a fixed-order wave sum) and diag_finish adds a row's pieces in
// order.
constexpr u64 kDiagPieceOwners = 4096;

__global__ void __launch_bounds__(64)
diag_pieces(const SegPiece *__restrict__ pieces, const u32 *__restrict__ w_leaf,
            const double *__restrict__ w_val, const unsigned char *__restrict__ kind,
            const double *__restrict__ mult, double *__restrict__ part) {
  const SegPiece pc = pieces[blockIdx.x];
  double acc = 0.0;
  for (u64 q = pc.o0 + threadIdx.x; q < pc.o1; q += 64) {
    const u32 leaf = w_leaf[q];
    if (kind[leaf] != 0 || (q > pc.first && w_leaf[q - 1] == leaf)) continue;
    const double v = w_val[q];
    acc += mult[leaf] * (v * v);
  }
  acc = wave_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void diag_finish(const SegOut *__restrict__ segs, u64 n,
                            const double *__restrict__ part, double *__restrict__ D) {
  const u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const SegOut so = segs[r];
  double acc = 0.0;
  for (u64 p = so.p0; p < so.p1; ++p) acc += part[p];
  D[so.row] = acc;
}

// Q[:, j0 : j0 + w] = W (W: n x w, Q: n x ldq, row-major)
__global__ void put_cols(const double *__restrict__ W, u64 n, int w, double *__restrict__ Q,
                         int ldq, int j0) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * w) return;
  Q[(t / w) * ldq + j0 + (int)(t % w)] = W[t];
}

__global__ void sub_inplace(double *__restrict__ W, const double *__restrict__ X, u64 n) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) W[t] -= X[t];
}

// out[i, c] = AX[i, c] - theta[c] X[i, c] for c < w (AX, X: n x ld; out n x w)
__global__ void ritz_residual(const double *__restrict__ AX, const double *__restrict__ X,
                              const double *__restrict__ theta, u64 n, int ld, int w,
                              double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * w) return;
  const u64 i = t / w;
  const int c = (int)(t % w);
  out[t] = AX[i * ld + c] - theta[c] * X[i * ld + c];
}

// M[r, p] = S[p, r] for p < j, 0 for j <= p < ldm (S: j x w row-major; M: w x ldm)
__global__ void coeff_rows(const double *__restrict__ S, int j, int w, int ldm,
                           double *__restrict__ M) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= w * ldm) return;
  const int r = t / ldm, p = t % ldm;
  M[t] = p < j ? S[(u64)p * w + r] : 0.0;
}

// columns c0 .. w-1 of W (n x w): fixed pseudo-random values in [-1, 1)
__global__ void fill_random_cols(double *__restrict__ W, u64 n, int w, int c0, u64 seed) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = w - c0;
  if (t >= n * nc) return;
  const u64 i = t / nc;
  const int c = c0 + (int)(t % nc);
  u64 z = seed + (i * w + c + 1) * 0x9E3779B97F4A7C15ull;  // splitmix64
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  W[i * w + c] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

}  // namespace

void hoevd_gram(Ctx &c, double *A) {
  if (!c.built) throw Error(SP_ERR_STATE, "tree not built");
  const u32 n = c.dim;
  const int m = (int)c.order - 1;
  if (m < 1) throw Error(SP_ERR_UNSUPPORTED, "HOEVD needs order >= 2");
  dim3 grid(n, (n + GRAM_TILE - 1) / GRAM_TILE);
  ScopedTimer t(c, "hoevd_gram");
  hoevd_gram_rows<<<grid, 64, 0, c.stream>>>(
      c.w_segptr.as<u64>(), c.w_leaf.as<u32>(), c.w_val.as<double>(), c.w_tuple.as<u32>(),
      m, c.levels[c.depth - 1].level_ptr.as<u64>(), c.own_rem.as<u32>(),
      c.own_val.as<double>(), n, A);
  SP_CHECK_LAUNCH();
}

namespace {

// The implicit HOEVD operator: A V for n x w row-major V (w a multiple of
// 16), from the leaf level's owner lists and the rem-sorted work list.  The
// set-up (leaf kinds, the diagonal, the compacted lists) runs on the device;
// only the pieces of the leaves and rows longer than a piece are cut on the
// host, from their owner ranges.
class HoevdOperator {
 public:
  explicit HoevdOperator(Ctx &c) : c_(c), n_(c.dim) {
    const DevLevel &lv = c.levels[c.depth - 1];
    L_ = lv.size;
    const u64 E = c.h_segptr.size() == n_ + 1 ? c.h_segptr[n_] : 0;
    if (c.h_segptr.size() != n_ + 1) throw Error(SP_ERR_STATE, "work list not built");
    const int m = (int)c.order - 1;
    const u64 *ptr = lv.level_ptr.as<u64>();
    mult_.alloc(std::max<u64>(L_, 1) * sizeof(double));
    C_.alloc(std::max<u64>(L_, 1) * kCols * sizeof(double));
    diag_.alloc(std::max<u64>(n_, 1) * sizeof(double));
    std::vector<SegPiece> lp, rp;
    std::vector<SegOut> ls, rs;
    {
      DevBuf kind(L_ + 1), light((L_ + 1) * sizeof(u64)), heavy((L_ + 1) * sizeof(u64));
      DevBuf lpos((L_ + 1) * sizeof(u64)), hpos((L_ + 1) * sizeof(u64));
      if (c.key_words == 1)
        leaf_mult<<<grid_for(L_, 256), 256, 0, c.stream>>>(lv.keys.as<u64>(), L_, m, c.bits,
                                                          mult_.as<double>());
      else
        leaf_mult<<<grid_for(L_, 256), 256, 0, c.stream>>>(lv.keys.as<u128>(), L_, m, c.bits,
                                                          mult_.as<double>());
      SP_CHECK_LAUNCH();
      leaf_kinds<<<grid_for(L_ + 1, 256), 256, 0, c.stream>>>(
          ptr, L_, c.own_rem.as<u32>(), kind.as<unsigned char>(), light.as<u64>(),
          heavy.as<u64>());
      SP_CHECK_LAUNCH();
      exclusive_scan_u64(light.as<u64>(), lpos.as<u64>(), L_ + 1, c.stream);
      exclusive_scan_u64(heavy.as<u64>(), hpos.as<u64>(), L_ + 1, c.stream);
      u64 nh = 0;
      SP_HIP(hipMemcpyAsync(&n_light_, lpos.as<u64>() + L_, sizeof(u64), hipMemcpyDeviceToHost,
                            c.stream));
      SP_HIP(hipMemcpyAsync(&nh, hpos.as<u64>() + L_, sizeof(u64), hipMemcpyDeviceToHost,
                            c.stream));
      SP_HIP(hipStreamSynchronize(c.stream));
      light_.alloc(std::max<u64>(n_light_, 1) * sizeof(u32));
      DevBuf hl(std::max<u64>(nh, 1) * sizeof(u32)), hr(std::max<u64>(nh, 1) * 2 * sizeof(u64));
      compact_ids<<<grid_for(L_, 256), 256, 0, c.stream>>>(light.as<u64>(), lpos.as<u64>(), L_,
                                                          light_.as<u32>());
      SP_CHECK_LAUNCH();
      compact_ids<<<grid_for(L_, 256), 256, 0, c.stream>>>(heavy.as<u64>(), hpos.as<u64>(), L_,
                                                          hl.as<u32>());
      SP_CHECK_LAUNCH();
      // pieces of the heavy leaves
      if (nh > 0) {
        leaf_ranges<<<grid_for(nh, 256), 256, 0, c.stream>>>(hl.as<u32>(), nh, ptr,
                                                            hr.as<u64>());
        SP_CHECK_LAUNCH();
        std::vector<u32> ids(nh);
        std::vector<u64> rg(2 * nh);
        SP_HIP(hipMemcpyAsync(ids.data(), hl.p, nh * sizeof(u32), hipMemcpyDeviceToHost,
                              c.stream));
        SP_HIP(hipMemcpyAsync(rg.data(), hr.p, 2 * nh * sizeof(u64), hipMemcpyDeviceToHost,
                              c.stream));
        SP_HIP(hipStreamSynchronize(c.stream));
        for (u64 i = 0; i < nh; ++i) cut(rg[2 * i], rg[2 * i + 1], ids[i], lp, ls);
      }
      // the diagonal
      {
        std::vector<SegPiece> dp;
        std::vector<SegOut> ds;
        for (u64 r = 0; r < n_; ++r)
          cut(c.h_segptr[r], c.h_segptr[r + 1], r, dp, ds, kDiagPieceOwners);
        DevBuf dpd, dsd, dpart(std::max<size_t>(dp.size(), 1) * sizeof(double));
        upload(dp, dpd);
        upload(ds, dsd);
        if (!dp.empty()) {
          diag_pieces<<<(unsigned)dp.size(), 64, 0, c.stream>>>(
              dpd.as<SegPiece>(), c.w_leaf.as<u32>(), c.w_val.as<double>(),
              kind.as<unsigned char>(), mult_.as<double>(), dpart.as<double>());
          SP_CHECK_LAUNCH();
        }
        diag_finish<<<grid_for(n_, 256), 256, 0, c.stream>>>(dsd.as<SegOut>(), n_,
                                                             dpart.as<double>(),
                                                             diag_.as<double>());
        SP_CHECK_LAUNCH();
        SP_HIP(hipStreamSynchronize(c.stream));  // dpd, dsd, dpart are freed here
      }
      // the work list without the single-owner leaves
      DevBuf ef((E + 1) * sizeof(u64)), epos((E + 1) * sizeof(u64));
      DevBuf seg2((n_ + 1) * sizeof(u64));
      entry_flags<<<grid_for(E + 1, 256), 256, 0, c.stream>>>(
          c.w_leaf.as<u32>(), E, kind.as<unsigned char>(), ef.as<u64>());
      SP_CHECK_LAUNCH();
      exclusive_scan_u64(ef.as<u64>(), epos.as<u64>(), E + 1, c.stream);
      u64 E2 = 0;
      SP_HIP(hipMemcpyAsync(&E2, epos.as<u64>() + E, sizeof(u64), hipMemcpyDeviceToHost,
                            c.stream));
      SP_HIP(hipStreamSynchronize(c.stream));
      leaf2_.alloc(std::max<u64>(E2, 1) * sizeof(u32));
      val2_.alloc(std::max<u64>(E2, 1) * sizeof(double));
      compact_entries<<<grid_for(E, 256), 256, 0, c.stream>>>(
          ef.as<u64>(), epos.as<u64>(), E, c.w_leaf.as<u32>(), c.w_val.as<double>(),
          leaf2_.as<u32>(), val2_.as<double>());
      SP_CHECK_LAUNCH();
      remap_rows<<<grid_for(n_ + 1, 256), 256, 0, c.stream>>>(c.w_segptr.as<u64>(),
                                                             epos.as<u64>(), n_, seg2.as<u64>());
      SP_CHECK_LAUNCH();
      std::vector<u64> hs(n_ + 1);
      SP_HIP(hipMemcpyAsync(hs.data(), seg2.p, (n_ + 1) * sizeof(u64), hipMemcpyDeviceToHost,
                            c.stream));
      SP_HIP(hipStreamSynchronize(c.stream));
      // pieces of every row (an empty row gets an empty piece range)
      for (u64 r = 0; r < n_; ++r) cut(hs[r], hs[r + 1], r, rp, rs);
    }
    n_leaf_pieces_ = lp.size();
    n_leaf_segs_ = ls.size();
    n_row_pieces_ = rp.size();
    upload(lp, leaf_pieces_);
    upload(ls, leaf_segs_);
    upload(rp, row_pieces_);
    upload(rs, row_segs_);
    part_.alloc(std::max<u64>(std::max(n_leaf_pieces_, n_row_pieces_), 1) * kCols *
                sizeof(double));
  }

  void apply(const double *V, int w, double *Y) {
    ScopedTimer t(c_, "hoevd_apply");
    const DevLevel &lv = c_.levels[c_.depth - 1];
    for (int col0 = 0; col0 < w; col0 += kCols) {
      if (n_light_ > 0) {
        leaf_rows_light<<<grid_for(n_light_ * kCols, 256), 256, 0, c_.stream>>>(
            light_.as<u32>(), n_light_, lv.level_ptr.as<u64>(), c_.own_rem.as<u32>(),
            c_.own_val.as<double>(), mult_.as<double>(), V, w, col0, C_.as<double>());
        SP_CHECK_LAUNCH();
      }
      if (n_leaf_pieces_ > 0) {
        seg_pieces<<<(unsigned)n_leaf_pieces_, 64, 0, c_.stream>>>(
            leaf_pieces_.as<SegPiece>(), c_.own_rem.as<u32>(), 1u, c_.own_val.as<double>(), V,
            w, col0, part_.as<double>());
        SP_CHECK_LAUNCH();
        seg_finish<<<grid_for(n_leaf_segs_ * kCols, 256), 256, 0, c_.stream>>>(
            leaf_segs_.as<SegOut>(), n_leaf_segs_, part_.as<double>(), mult_.as<double>(),
            nullptr, nullptr, C_.as<double>(), kCols, 0);
        SP_CHECK_LAUNCH();
      }
      if (n_row_pieces_ > 0) {
        seg_pieces<<<(unsigned)n_row_pieces_, 64, 0, c_.stream>>>(
            row_pieces_.as<SegPiece>(), leaf2_.as<u32>(), 0u, val2_.as<double>(),
            C_.as<double>(), kCols, 0, part_.as<double>());
        SP_CHECK_LAUNCH();
      }
      seg_finish<<<grid_for(n_ * kCols, 256), 256, 0, c_.stream>>>(
          row_segs_.as<SegOut>(), n_, part_.as<double>(), nullptr, diag_.as<double>(), V, Y, w,
          col0);
      SP_CHECK_LAUNCH();
    }
  }

 private:
  static void cut(u64 o0, u64 o1, u64 row, std::vector<SegPiece> &pieces,
                  std::vector<SegOut> &segs, u64 len = kPieceOwners) {
    const u64 p0 = pieces.size();
    for (u64 q = o0; q < o1; q += len) pieces.push_back({q, std::min(o1, q + len), o0});
    segs.push_back({row, p0, (u64)pieces.size()});
  }
  template <class T>
  void upload(const std::vector<T> &h, DevBuf &d) {
    d.alloc(std::max<size_t>(h.size(), 1) * sizeof(T));
    if (!h.empty())
      SP_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice,
                            c_.stream));
    SP_HIP(hipStreamSynchronize(c_.stream));  // h is freed by the caller
  }

  Ctx &c_;
  u64 n_, L_ = 0, n_light_ = 0;
  u64 n_leaf_pieces_ = 0, n_leaf_segs_ = 0, n_row_pieces_ = 0;
  DevBuf mult_, C_, diag_, light_, leaf2_, val2_, part_;
  DevBuf leaf_pieces_, leaf_segs_, row_pieces_, row_segs_;
};

// Block width and basis size of the Krylov solve; false where it does not
// apply (a basis of at least four blocks must fit in dim, 2048 columns and
// 8 GB for the basis and its image).
bool krylov_shape(u64 n, u32 R, int &b, int &kmax) {
  b = (int)((R + kCols - 1) / kCols * kCols);
  u64 cap = std::min<u64>({n, 32 * (u64)b, 2048, (u64)(5.0e8 / (double)n)});
  kmax = (int)(cap / b * b);
  return kmax >= 4 * b;
}

constexpr u64 kDenseMaxDim = 4096;  // dense HOEVD up to this dim by default
constexpr double kRitzTol = 1e-12;  // residual ||A x - theta x|| <= kRitzTol * theta_max
constexpr int kMaxRestarts = 40;

// Top-R eigenpairs of A = X_(1) X_(1)^T by block Lanczos on the implicit
// operator, with full reorthogonalisation (two projections per block), an
// explicit Rayleigh-Ritz step on B = Q^T A Q every fourth block, and thick
// restart with the leading half of the Ritz vectors when the basis is full.
// A block is orthonormalised by SVQB (Stathopoulos & Wu, SIAM J. Sci.
// Comput. 23 (2002)): the eigen-decomposition of its b x b Gram, dropping
// directions that are numerically inside the basis and replacing them by
// fixed pseudo-random ones.  Stops when every wanted Ritz pair has residual
// norm at most kRitzTol times the largest Ritz value.  Output as the dense
// path: U_out's columns and eig_out ascending.
void hoevd_krylov(Ctx &c, u32 R, double *U_out, double *eig_out) {
  const u64 n = c.dim;
  int b = 0, kmax = 0;
  if (!krylov_shape(n, R, b, kmax))
    throw Error(SP_ERR_UNSUPPORTED, "rank too large for the dimension in the Krylov HOEVD");
  const int keep = std::max(b, kmax / b / 2 * b);
  HoevdOperator op(c);
  const size_t nb8 = n * (size_t)b * sizeof(double);
  DevBuf Q(n * (size_t)kmax * sizeof(double)), AQ(n * (size_t)kmax * sizeof(double));
  DevBuf X(nb8), W(nb8), T(nb8);
  DevBuf Rv(n * (size_t)keep * sizeof(double)), ARv(n * (size_t)keep * sizeof(double));
  DevBuf H((size_t)b * kmax * sizeof(double)), M((size_t)keep * kmax * sizeof(double));
  DevBuf Bd((size_t)kmax * kmax * sizeof(double)), S((size_t)kmax * keep * sizeof(double));
  DevBuf th((size_t)keep * sizeof(double)), G((size_t)std::max<u32>(b, R) * b * sizeof(double));
  SP_HIP(hipMemsetAsync(Q.p, 0, Q.bytes, c.stream));
  SP_HIP(hipMemsetAsync(AQ.p, 0, AQ.bytes, c.stream));
  const dim3 g_nb(grid_for(n * b, 256));
  u64 seed = 0x484f455644ull;

  auto download = [&](const DevBuf &d, std::vector<double> &h, size_t count) {
    h.resize(count);
    SP_HIP(hipMemcpyAsync(h.data(), d.p, count * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
  };
  auto upload = [&](const std::vector<double> &h, DevBuf &d) {
    SP_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          c.stream));
  };
  // V -= Q (V^T Q)^T
  auto project = [&](DevBuf &V) {
    dense_atb(c, V.as<double>(), b, Q.as<double>(), kmax, n, H.as<double>());
    dense_ymt(c, Q.as<double>(), n, kmax, H.as<double>(), b, T.as<double>());
    sub_inplace<<<grid_for(n * b, 256), 256, 0, c.stream>>>(V.as<double>(), T.as<double>(),
                                                            n * b);
    SP_CHECK_LAUNCH();
  };
  // one SVQB pass: V <- V Vg diag(w)^(-1/2) on the directions with w > floor,
  // zero columns for the others (last); returns the number kept
  std::vector<double> hg, wv, vv, hm;
  auto svqb = [&](DevBuf &V, double floor) {
    dense_atb(c, V.as<double>(), b, V.as<double>(), b, n, G.as<double>());
    download(G, hg, (size_t)b * b);
    for (int i = 0; i < b; ++i)
      for (int k = 0; k < i; ++k) hg[(size_t)i * b + k] = hg[(size_t)k * b + i];
    dense_small_eig(hg, b, wv, vv);
    int kept = 0;
    hm.assign((size_t)b * b, 0.0);
    for (int r = 0; r < b; ++r) {
      if (!(wv[r] > floor)) continue;
      const double sc = 1.0 / std::sqrt(wv[r]);
      for (int p = 0; p < b; ++p) hm[(size_t)kept * b + p] = vv[(size_t)p * b + r] * sc;
      ++kept;
    }
    upload(hm, H);
    dense_ymt(c, V.as<double>(), n, b, H.as<double>(), b, T.as<double>());
    SP_HIP(hipMemcpyAsync(V.p, T.p, nb8, hipMemcpyDeviceToDevice, c.stream));
    return kept;
  };
  // V <- an orthonormal block orthogonal to the basis, spanning V's part
  // outside it (ref: the scale of V, below which a remainder is rounding)
  auto orth_block = [&](DevBuf &V, double ref) {
    for (int attempt = 0; attempt < 4; ++attempt) {
      project(V);
      svqb(V, (kRitzTol * ref) * (kRitzTol * ref));
      project(V);
      const int kept = svqb(V, 0.5);
      if (kept == b) return;
      fill_random_cols<<<g_nb, 256, 0, c.stream>>>(V.as<double>(), n, b, kept, seed++);
      SP_CHECK_LAUNCH();
      ref = 1.0;
    }
    throw Error(SP_ERR_NUMERIC, "Krylov HOEVD: no block orthogonal to the basis");
  };

  fill_random_cols<<<g_nb, 256, 0, c.stream>>>(X.as<double>(), n, b, 0, seed++);
  SP_CHECK_LAUNCH();
  orth_block(X, 1.0);
  std::vector<double> B((size_t)kmax * kmax, 0.0), hh, hb, theta, hres;
  double theta_max = 0.0;
  int j = 0, restarts = 0;
  ScopedTimer tk(c, "hoevd_krylov");
  for (;;) {
    put_cols<<<g_nb, 256, 0, c.stream>>>(X.as<double>(), n, b, Q.as<double>(), kmax, j);
    SP_CHECK_LAUNCH();
    op.apply(X.as<double>(), b, W.as<double>());
    put_cols<<<g_nb, 256, 0, c.stream>>>(W.as<double>(), n, b, AQ.as<double>(), kmax, j);
    SP_CHECK_LAUNCH();
    // B[:, j : j+b] = Q^T (A X), and its transpose
    dense_atb(c, W.as<double>(), b, Q.as<double>(), kmax, n, H.as<double>());
    download(H, hh, (size_t)b * kmax);
    for (int r = 0; r < b; ++r) {
      for (int p = 0; p < j; ++p)
        B[(size_t)p * kmax + j + r] = B[(size_t)(j + r) * kmax + p] = hh[(size_t)r * kmax + p];
      for (int s2 = 0; s2 < b; ++s2)
        B[(size_t)(j + r) * kmax + j + s2] =
            0.5 * (hh[(size_t)r * kmax + j + s2] + hh[(size_t)s2 * kmax + j + r]);
      theta_max = std::max(theta_max, std::fabs(B[(size_t)(j + r) * kmax + j + r]));
    }
    j += b;
    const bool full = j + b > kmax;
    if (full || (j / b) % 4 == 0) {
      // Rayleigh-Ritz on B[:j, :j]
      const int nw = full ? keep : (int)R;
      hb.resize((size_t)j * j);
      for (int p = 0; p < j; ++p)
        std::copy(B.begin() + (size_t)p * kmax, B.begin() + (size_t)p * kmax + j,
                  hb.begin() + (size_t)p * j);
      upload(hb, Bd);
      eig_top(c, Bd.as<double>(), (u64)j, (u32)nw, S.as<double>(), theta);
      coeff_rows<<<grid_for((u64)nw * kmax, 256), 256, 0, c.stream>>>(S.as<double>(), j, nw,
                                                                      kmax, M.as<double>());
      SP_CHECK_LAUNCH();
      dense_ymt(c, Q.as<double>(), n, kmax, M.as<double>(), nw, Rv.as<double>());
      dense_ymt(c, AQ.as<double>(), n, kmax, M.as<double>(), nw, ARv.as<double>());
      upload(theta, th);
      theta_max = std::max(theta_max, std::fabs(theta[0]));
      // residuals of the wanted pairs
      ritz_residual<<<grid_for(n * R, 256), 256, 0, c.stream>>>(
          ARv.as<double>(), Rv.as<double>(), th.as<double>(), n, nw, (int)R, T.as<double>());
      SP_CHECK_LAUNCH();
      dense_atb(c, T.as<double>(), R, T.as<double>(), R, n, G.as<double>());
      download(G, hres, (size_t)R * R);
      double worst = 0.0;
      for (u32 i = 0; i < R; ++i)
        worst = std::max(worst, std::sqrt(std::fabs(hres[(size_t)i * R + i])));
      if (!std::isfinite(worst)) throw Error(SP_ERR_NUMERIC, "Krylov HOEVD: non-finite residual");
      if (worst <= kRitzTol * theta_max || (u64)j >= n) {
        std::vector<double> hv;
        download(Rv, hv, n * (size_t)nw);
        for (u64 i = 0; i < n; ++i)
          for (u32 k = 0; k < R; ++k) U_out[i * R + k] = hv[i * nw + (R - 1 - k)];
        if (eig_out)
          for (u32 k = 0; k < R; ++k) eig_out[k] = theta[R - 1 - k];
        return;
      }
      if (full) {
        if (++restarts > kMaxRestarts)
          throw Error(SP_ERR_NUMERIC,
                      "Krylov HOEVD did not converge (SP_HOEVD=dense selects the dense solver)");
        // thick restart: keep the leading Ritz vectors and their images
        SP_HIP(hipMemsetAsync(Q.p, 0, Q.bytes, c.stream));
        SP_HIP(hipMemsetAsync(AQ.p, 0, AQ.bytes, c.stream));
        put_cols<<<grid_for(n * keep, 256), 256, 0, c.stream>>>(Rv.as<double>(), n, keep,
                                                                Q.as<double>(), kmax, 0);
        SP_CHECK_LAUNCH();
        put_cols<<<grid_for(n * keep, 256), 256, 0, c.stream>>>(ARv.as<double>(), n, keep,
                                                                AQ.as<double>(), kmax, 0);
        SP_CHECK_LAUNCH();
        std::fill(B.begin(), B.end(), 0.0);
        for (int i = 0; i < keep; ++i) B[(size_t)i * kmax + i] = theta[i];
        j = keep;
        // next block: the residuals of the leading b Ritz vectors
        ritz_residual<<<g_nb, 256, 0, c.stream>>>(ARv.as<double>(), Rv.as<double>(),
                                                  th.as<double>(), n, keep, b, X.as<double>());
        SP_CHECK_LAUNCH();
        orth_block(X, theta_max);
        continue;
      }
    }
    SP_HIP(hipMemcpyAsync(X.p, W.p, nb8, hipMemcpyDeviceToDevice, c.stream));
    orth_block(X, theta_max);
  }
}

}  // namespace

void hoevd_init(Ctx &c, u32 R, double *U_out, double *eig_out) {
  const u64 n = c.dim;
  if (R < 1 || R > n) throw Error(SP_ERR_ARG, "rank must be in [1, dim]");
  if (!c.built) throw Error(SP_ERR_STATE, "tree not built");
  int b = 0, kmax = 0;
  if (c.hoevd_method == 2 ||
      (c.hoevd_method == 0 && n > kDenseMaxDim && krylov_shape(n, R, b, kmax))) {
    hoevd_krylov(c, R, U_out, eig_out);
    return;
  }
  if (n > 0x7fffffffull / n) throw Error(SP_ERR_UNSUPPORTED, "dim too large for dense HOEVD");
  DevBuf A(n * n * sizeof(double));
  hoevd_gram(c, A.as<double>());
  if (c.eig_method >= 2) {
    // only the R wanted pairs (eig_top: tridiagonal reduction, host bisection
    // and inverse iteration for R pairs, back-transform of R vectors); the
    // columns come out descending, matrix::syevd's last R columns ascending
    DevBuf V(n * R * sizeof(double));
    std::vector<double> theta;
    {
      ScopedTimer t(c, "hoevd_eig");
      eig_top(c, A.as<double>(), n, R, V.as<double>(), theta);
    }
    std::vector<double> Vh(n * R);
    SP_HIP(hipMemcpyAsync(Vh.data(), V.p, n * R * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
    for (u64 i = 0; i < n; ++i)
      for (u32 j = 0; j < R; ++j) U_out[i * R + j] = Vh[i * R + (R - 1 - j)];
    if (eig_out)
      for (u32 j = 0; j < R; ++j) eig_out[j] = theta[R - 1 - j];
    return;
  }
  // full divide-and-conquer spectrum, as matrix::syevd (LAPACKE_dsyevd), on
  // A scaled into range by a power of two (LAPACK's dsyevd scales its input)
  const int ex = scale_to_unit(c, A.as<double>(), n * n);
  DevBuf W(2 * n * sizeof(double) + sizeof(rocblas_int));
  double *D = W.as<double>();
  double *E = D + n;
  rocblas_int *info = reinterpret_cast<rocblas_int *>(E + n);
  rocblas_handle h = blas_handle(c);
  {
    ScopedTimer t(c, "hoevd_syevd");
    const rocblas_status st =
        rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
                         A.as<double>(), (rocblas_int)n, D, E, info);
    if (st != rocblas_status_success) throw Error(SP_ERR_DEVICE, "rocsolver_dsyevd failed");
  }
  rocblas_int hinfo = 0;
  SP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  if (hinfo != 0) throw Error(SP_ERR_NUMERIC, "HOEVD eigensolver did not converge");
  // U[i][j] = eigenvector n-R+j (ascending order), matrixops.cpp:92-97
  DevBuf Ur(n * R * sizeof(double));
  transpose_cols<<<grid_for(n * R, 256), 256, 0, c.stream>>>(
      A.as<double>() + (n - R) * n, n, R, Ur.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(U_out, Ur.p, n * R * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  if (eig_out)
    SP_HIP(hipMemcpyAsync(eig_out, D + (n - R), R * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  if (eig_out)
    for (u32 j = 0; j < R; ++j) eig_out[j] = std::ldexp(eig_out[j], -ex);
}

}  // namespace spamd
