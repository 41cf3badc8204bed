// Dense Y-side operations of the Tucker loops, FP64 on the device.
//
//   gemmtn   core = U^T Y                     (src/dense/matrixops.cpp:45-56)
//   symgemm  U = Y diag(W) core^T             (src/dense/matrixops.cpp:102-124)
//   gram     G = diag(sqrt W) Y^T Y diag(sqrt W), the P x P Gram whose
//            nonzero spectrum equals that of symsyrk's n x n Y diag(W) Y^T
//            (src/dense/matrixops.cpp:126-146)
//   leftsvd  top-R left singular vectors of the unpacked Y: with
//            Z = Y diag(sqrt W), Z Z^T = Y_unpacked Y_unpacked^T, so
//            U = Z V_R Lambda_R^{-1/2} from the top-R eigenpairs of G
//   qr       thin Q of U by Cholesky-QR2, rocSOLVER geqrf/orgqr when U is
//            ill-conditioned           (src/dense/matrixops.cpp:58-75)
//   regularize  row clip to delta + left singular vectors of U
//            (src/sparse/spsymtensor.cpp:168-199)
// Reductions over the n rows go through fixed-order partial sums, so results
// are deterministic.
#include "sp_internal.hpp"
#include "wave.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <limits>

namespace spamd {

namespace {

// Relative eigenvalue floor of a Gram matrix below which a direction counts
// as numerically absent (singular value ratio ~3e-6).
constexpr double kGramRankTol = 1e-11;

typedef double d4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64 operand layout: lane l holds A[m = l%16][k = l/16]
// and B[k = l/16][n = l%16]; the accumulator's register q holds
// C[m = 4q + l/16][n = l%16].
__device__ inline d4 mfma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// partial[b][r][p] = sum over rows i of block b of A[i][r] * B[i][p]
// (A: n x R, B: n x P, both row-major).  FP64 MFMA: a wavefront owns 16
// columns of B and one 16-wide block of r, and walks its rows four at a time
// (K = rows); two accumulators alternate to keep the matrix pipe busy.
// grid (row blocks, ceil(P/64)); 4 wavefronts per block.
__global__ void __launch_bounds__(256)
atb_partial(const double *__restrict__ A, u32 R, const double *__restrict__ B,
            u64 P, u64 n, u64 rows_per, double *__restrict__ partial) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kr = lane >> 4, kc = lane & 15;
  const u64 col = (u64)blockIdx.y * 64 + wave * 16 + kc;
  const u64 i0 = (u64)blockIdx.x * rows_per;
  const u64 i1 = min(n, i0 + rows_per);
  const bool cin = col < P;
  for (u32 r0 = 0; r0 < R; r0 += 16) {
    const bool rin = r0 + kc < R;
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    u64 i = i0;
    for (; i + 8 <= i1; i += 8) {
      const u64 ia = i + kr, ib = i + 4 + kr;
      const double a0 = rin ? A[ia * R + r0 + kc] : 0.0;
      const double b0 = cin ? B[ia * P + col] : 0.0;
      const double a1 = rin ? A[ib * R + r0 + kc] : 0.0;
      const double b1 = cin ? B[ib * P + col] : 0.0;
      acc0 = mfma64(a0, b0, acc0);
      acc1 = mfma64(a1, b1, acc1);
    }
    for (; i < i1; i += 4) {
      const u64 ia = i + kr;
      const bool in = ia < i1;
      const double a0 = (rin && in) ? A[ia * R + r0 + kc] : 0.0;
      const double b0 = (cin && in) ? B[ia * P + col] : 0.0;
      acc0 = mfma64(a0, b0, acc0);
    }
    // C[m = 4q + kr][n = kc]: m is r - r0, n the column
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u32 r = r0 + 4 * q + kr;
      if (r < R && cin) partial[((u64)blockIdx.x * R + r) * P + col] = acc0[q] + acc1[q];
    }
  }
}

// (((0 + x_0) + x_1) + ...) + x_{count-1} with x_i = p[i * stride]: the same
// additions in the same order as the plain loop (so the same bits), with the
// loads of eight terms issued before their additions rather than one load
// wait per term.
__device__ inline double sum_strided(const double *__restrict__ p, u64 stride, u64 count) {
  double s = 0.0;
  u64 i = 0;
  for (; i + 8 <= count; i += 8) {
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = p[(i + k) * stride];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += x[k];
  }
  for (; i < count; ++i) s += p[i * stride];
  return s;
}

__global__ void sum_partials(const double *__restrict__ partial, u64 nb, u64 len,
                             double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  out[t] = sum_strided(partial + t, len, nb);
}

// The same sums for mid-size outputs: a block takes 64 consecutive entries
// with 4 threads each (thread q adds partials q, q+4, ..., loads coalesced
// across the entries), then the 4 sums of an entry in a fixed order.
__global__ void __launch_bounds__(256) sum_partials_cols(const double *__restrict__ partial,
                                                         u64 nb, u64 len,
                                                         double *__restrict__ out) {
  __shared__ double acc[4][64];
  const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
  const u64 t = (u64)blockIdx.x * 64 + e;
  double s = 0.0;
  if (t < len && (u64)q < nb) s = sum_strided(partial + (u64)q * len + t, 4 * len, (nb - q + 3) / 4);
  acc[q][e] = s;
  __syncthreads();
  if (q == 0 && t < len) out[t] = (acc[0][e] + acc[1][e]) + (acc[2][e] + acc[3][e]);
}

// Largest |a_i| of each block's grid-stride share (max is exact in any order).
__global__ void __launch_bounds__(256) absmax_partial(const double *__restrict__ a, u64 n,
                                                      double *__restrict__ part) {
  __shared__ double m[256];
  double v = 0.0;
  for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < n; i += (u64)gridDim.x * 256)
    v = fmax(v, fabs(a[i]));
  m[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = m[0];
}

__global__ void scale_by(double *__restrict__ a, u64 n, double s) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] *= s;
}

// The largest of the nb partial maxima (fixed-order tree), and the power of
// two that brings it into [1, 2) when it lies outside [2^-400, 2^400]
// (finite, nonzero; 0 otherwise) into *ex.  One block of 256 threads.
__global__ void __launch_bounds__(256) scale_pick(const double *__restrict__ part, int nb,
                                                  int *__restrict__ ex) {
  __shared__ double m[256];
  m[threadIdx.x] = (int)threadIdx.x < nb ? part[threadIdx.x] : 0.0;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a = m[0];
    *ex = a > 0.0 && isfinite(a) && (a < 0x1p-400 || a > 0x1p400) ? -ilogb(a) : 0;
  }
}

__global__ void scale_by_pow2(double *__restrict__ a, u64 n, const int *__restrict__ ex) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const int e = *ex;
  if (i < n && e != 0) a[i] = ldexp(a[i], e);
}

// The same sums for short outputs (a few hundred entries, where one thread
// per entry would add its nb partials one after another): one wavefront per
// entry, lanes over the partials, then a fixed-order wave sum.
__global__ void __launch_bounds__(256) sum_partials_wave(const double *__restrict__ partial,
                                                         u64 nb, u64 len,
                                                         double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * 4 + (threadIdx.x >> 6);
  const u64 lane = threadIdx.x & 63;
  if (t >= len) return;  // whole wavefronts leave together
  double s = 0.0;
  for (u64 b = lane; b < nb; b += 64) s += partial[b * len + t];
  s = wave_sum(s);
  if (lane == 0) out[t] = s;
}

// out[i][r] = sum_p Y[i][p] * M[r][p]  (Y: n x P, M: R x P, out: n x R).
// FP64 MFMA with K = P: a block stages a 64-row x 32-column slab of Y and the
// matching 16 x 32 slab of M in LDS (coalesced row reads), each wavefront
// multiplies its 16 rows.  grid (ceil(n/64), ceil(R/16)).
constexpr int YK = 32;
__global__ void __launch_bounds__(256)
y_times_mt(const double *__restrict__ Y, u64 n, u64 P, const double *__restrict__ M,
           u32 R, double *__restrict__ out) {
  __shared__ double sy[64][YK + 1];
  __shared__ double sm[16][YK + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kr = lane >> 4, kc = lane & 15;
  const u64 row0 = (u64)blockIdx.x * 64;
  const u32 r0 = blockIdx.y * 16;
  d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  // the next slab is loaded into registers while the current one in LDS is
  // multiplied (thread t holds elements t + 256 j of each slab)
  constexpr int NY = 64 * YK / 256, NM = 16 * YK / 256;
  double ry[NY], rm[NM];
  auto fetch = [&](u64 k0) {
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      const int t = threadIdx.x + 256 * j, rr = t / YK, kk = t % YK;
      const u64 i = row0 + rr, k = k0 + kk;
      ry[j] = (i < n && k < P) ? Y[i * P + k] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NM; ++j) {
      const int t = threadIdx.x + 256 * j, rr = t / YK, kk = t % YK;
      const u64 k = k0 + kk;
      rm[j] = (r0 + rr < R && k < P) ? M[(u64)(r0 + rr) * P + k] : 0.0;
    }
  };
  fetch(0);
  for (u64 k0 = 0; k0 < P; k0 += YK) {
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      const int t = threadIdx.x + 256 * j;
      sy[t / YK][t % YK] = ry[j];
    }
#pragma unroll
    for (int j = 0; j < NM; ++j) {
      const int t = threadIdx.x + 256 * j;
      sm[t / YK][t % YK] = rm[j];
    }
    __syncthreads();
    if (k0 + YK < P) fetch(k0 + YK);
#pragma unroll
    for (int ks = 0; ks < YK / 4; ks += 2) {
      // A[m][k] = Y[row0 + 16 wave + m][k0 + 4 ks + k], B[k][nn] = M[r0 + nn][k0 + 4 ks + k]
      acc0 = mfma64(sy[wave * 16 + kc][4 * ks + kr], sm[kc][4 * ks + kr], acc0);
      acc1 = mfma64(sy[wave * 16 + kc][4 * ks + 4 + kr], sm[kc][4 * ks + 4 + kr], acc1);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const u64 i = row0 + wave * 16 + 4 * q + kr;
    const u32 r = r0 + kc;
    if (i < n && r < R) out[i * R + r] = acc0[q] + acc1[q];
  }
}

inline dim3 ymt_grid(u64 n, u32 R) {
  return dim3((unsigned)((n + 63) / 64), (R + 15) / 16);
}

__global__ void scale_cols(const double *__restrict__ core, const double *__restrict__ W,
                           u64 R, u64 P, double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R * P) out[t] = core[t] * W[t % P];
}

// Gram tile kernel: partial[kb][tile] for 64x64 tiles (tp <= tq) of Y^T Y.
constexpr int GT = 64;
constexpr int GK = 32;
constexpr int GL = GK * GT / 256;  // slab elements per thread
__global__ void __launch_bounds__(256, 4)
gram_partial(const double *__restrict__ Y, u64 n, u64 P, u64 rows_per, u32 ntiles, u32 nk,
             const int2 *__restrict__ tiles, double *__restrict__ partial) {
  // FP64 MFMA: wavefront w computes the 32 x 32 quadrant (w/2, w%2) of the
  // 64 x 64 tile as 2 x 2 blocks of v_mfma_f64_16x16x4f64 with K = rows.
  // The next GK-row slab is loaded into registers while the current one in
  // LDS is multiplied.
  //
  // Workgroups are dealt to the 8 XCDs round-robin by linear id, so block b
  // runs on XCD b % 8.  The work items (chunk, tile), chunk-major, are split
  // into 8 equal contiguous ranges and block 8 s + x takes item s of range x:
  // each XCD gets the same number of items, and the tiles of a row chunk run
  // on one XCD at about the same time, reading the chunk's rows of Y from
  // that XCD's L2 after the first reader.  Blocks past the end leave at once.
  __shared__ double sa[GK][GT + 1];
  __shared__ double sb[GK][GT + 1];
  const u32 items = ntiles * nk, per = (items + 7) / 8;
  const u32 slot = blockIdx.x >> 3, item = (blockIdx.x & 7) * per + slot;
  if (slot >= per || item >= items) return;  // the whole block, before any barrier
  const u32 tile = item % ntiles, chunk = item / ntiles;
  const int2 tl = tiles[tile];
  const u64 p0 = (u64)tl.x * GT, q0 = (u64)tl.y * GT;
  const bool diag = tl.x == tl.y;
  const u64 i0 = (u64)chunk * rows_per;
  const u64 i1 = min(n, i0 + rows_per);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kr = lane >> 4, kc = lane & 15;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  // this thread's slab elements: row kk = t / GT, column c = t % GT with
  // t = threadIdx.x + 256 j (c is the same for every j)
  const int c = threadIdx.x % GT, kk0 = threadIdx.x / GT;
  const bool ca = p0 + c < P, cb = q0 + c < P;
  double ra[GL], rb[GL];
  auto fetch = [&](u64 k) {
#pragma unroll
    for (int j = 0; j < GL; ++j) {
      const u64 i = k + kk0 + 4 * j;
      ra[j] = (i < i1 && ca) ? Y[i * P + p0 + c] : 0.0;
      rb[j] = (!diag && i < i1 && cb) ? Y[i * P + q0 + c] : 0.0;
    }
  };
  d4 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = d4{0.0, 0.0, 0.0, 0.0};
  if (i0 < i1) fetch(i0);
  for (u64 k = i0; k < i1; k += GK) {
#pragma unroll
    for (int j = 0; j < GL; ++j) {
      sa[kk0 + 4 * j][c] = ra[j];
      sb[kk0 + 4 * j][c] = rb[j];
    }
    __syncthreads();
    if (k + GK < i1) fetch(k + GK);
    double(*bs)[GT + 1] = diag ? sa : sb;
#pragma unroll
    for (int ks = 0; ks < GK / 4; ++ks) {
      const int kk = 4 * ks + kr;
      const double a0 = sa[kk][wr + kc], a1 = sa[kk][wr + 16 + kc];
      const double b0 = bs[kk][wc + kc], b1 = bs[kk][wc + 16 + kc];
      acc[0][0] = mfma64(a0, b0, acc[0][0]);
      acc[0][1] = mfma64(a0, b1, acc[0][1]);
      acc[1][0] = mfma64(a1, b0, acc[1][0]);
      acc[1][1] = mfma64(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  double *out = partial + ((u64)chunk * ntiles + tile) * GT * GT;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        out[(wr + 16 * u + 4 * q + kr) * GT + wc + 16 * v + kc] = acc[u][v][q];
}

__global__ void gram_finish(const double *__restrict__ partial, u64 nk, u64 ntiles,
                            const int2 *__restrict__ tiles, u64 P,
                            const double *__restrict__ W, double *__restrict__ G) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles * GT * GT) return;
  const u64 tile = t / (GT * GT);
  const int u = (int)((t % (GT * GT)) / GT), v = (int)(t % GT);
  const int2 tl = tiles[tile];
  const u64 p = (u64)tl.x * GT + u, q = (u64)tl.y * GT + v;
  if (p >= P || q >= P) return;
  if (tl.x == tl.y && u > v) return;  // each symmetric pair written once
  double s = sum_strided(partial + tile * GT * GT + u * GT + v, ntiles * GT * GT, nk);
  s *= sqrt(W[p]) * sqrt(W[q]);
  G[p * P + q] = s;
  G[q * P + p] = s;
}

// M[r][p] = sqrt(W[p]) * V[p][r] / sqrt(theta_r)  (V: P x R row-major, the
// top-R eigenvectors of G, theta descending)
__global__ void svd_map(const double *__restrict__ V, const double *__restrict__ theta,
                        const double *__restrict__ W, u64 P, u32 R, double floor_,
                        double *__restrict__ M) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (u64)R * P) return;
  const u64 r = t / P, p = t % P;
  const double l = theta[r] > floor_ ? theta[r] : floor_;
  M[t] = sqrt(W[p]) * V[p * R + r] / sqrt(l);
}

// V[p][r] = column P-1-r of the column-major eigenvector matrix Z (the
// eigenvector of the r-th largest eigenvalue from an ascending solver)
__global__ void top_columns(const double *__restrict__ Z, u64 P, u32 R,
                            double *__restrict__ V) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * R) return;
  const u64 p = t / R, r = t % R;
  V[t] = Z[(P - 1 - r) * P + p];
}

__global__ void transpose_nr(const double *__restrict__ in, u64 rows, u64 cols,
                             double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const u64 i = t / cols, j = t % cols;
  out[j * rows + i] = in[t];
}

// out[r][p] = C[r][p] * sqrt(W[p])  (C, out: R x P)
__global__ void scale_cols_sqrt(const double *__restrict__ C, const double *__restrict__ W,
                                u64 R, u64 P, double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R * P) out[t] = C[t] * sqrt(W[t % P]);
}

// res[r] = sum_p (GY[r][p] - theta[r] Y[r][p])^2: one 256-thread block per
// row r, strided partial sums, then a fixed-order tree.
__global__ void __launch_bounds__(256) resid_sq(const double *__restrict__ GY,
                                                const double *__restrict__ Y,
                                                const double *__restrict__ theta, u64 P,
                                                double *__restrict__ res) {
  __shared__ double s[256];
  const u64 r = blockIdx.x;
  const double th = theta[r];
  double a = 0.0;
  for (u64 p = threadIdx.x; p < P; p += 256) {
    const double d = GY[r * P + p] - th * Y[r * P + p];
    a += d * d;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[r] = s[0];
}

// out[r][p] = (GY[r][p] - theta[r] Y[r][p]) * inv[r]: Ritz residual
// directions, each scaled by the inverse of its norm
__global__ void ritz_dir(const double *__restrict__ GY, const double *__restrict__ Y,
                         const double *__restrict__ theta, const double *__restrict__ inv,
                         u64 R, u64 P, double *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R * P) return;
  const u64 r = t / P;
  out[t] = (GY[t] - theta[r] * Y[t]) * inv[r];
}

// Columns from..R-1 of U (n x R) <- fixed pseudo-random values in [-1, 1)
// (splitmix64 of the entry's position), the start of an orthonormal
// completion when the span to orthonormalise has fewer than R dimensions.
__global__ void fill_columns(double *__restrict__ U, u64 n, u32 R, u32 from) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u32 w = R - from;
  if (t >= n * w) return;
  const u64 i = t / w;
  const u32 j = from + (u32)(t % w);
  u64 z = i * (u64)R + j + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  U[i * R + j] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

__global__ void clip_rows(double *__restrict__ U, u64 n, u32 R, double delta) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (u32 j = 0; j < R; ++j) s += U[i * R + j] * U[i * R + j];
  const double nrm = sqrt(s);
  if (delta < nrm) {
    const double mul = delta / nrm;
    for (u32 j = 0; j < R; ++j) U[i * R + j] *= mul;
  }
}

// Host: cyclic Jacobi eigen-decomposition of a small symmetric matrix
// (row-major n x n).  Eigenvalues descending in w, eigenvectors in the
// columns of V (row-major).
void jacobi_eig(std::vector<double> A, int n, std::vector<double> &w,
                std::vector<double> &V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a = A[(size_t)i * n + j];
        tot += a * a;
        if (i != j) off += a * a;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) {
    return A[(size_t)a * n + a] > A[(size_t)b * n + b];
  });
  w.resize(n);
  std::vector<double> V2((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[(size_t)ord[j] * n + ord[j]];
    for (int i = 0; i < n; ++i) V2[(size_t)i * n + j] = V[(size_t)i * n + ord[j]];
  }
  V.swap(V2);
}

rocblas_handle handle_for(Ctx &c) {
  static thread_local rocblas_handle h = nullptr;
  if (!h) {
    if (rocblas_create_handle(&h) != rocblas_status_success)
      throw Error(SP_ERR_DEVICE, "rocblas_create_handle failed");
  }
  rocblas_set_stream(h, c.stream);
  return h;
}

// out = A^T B over n rows (A: n x R, B: n x P) -> R x P, deterministic.
void atb(Ctx &c, const double *A, u32 R, const double *B, u64 P, u64 n, double *out,
         const char *name) {
  const u64 ctiles = (P + 63) / 64;
  // about 2048 blocks in all but at most 256 row blocks (the partial sums
  // are added serially), at least 32 rows each, rows a multiple of 8
  u64 nb = std::max<u64>(1, std::min<u64>(std::min<u64>(2048 / ctiles + 1, 256),
                                          (n + 31) / 32));
  const u64 rows_per = ((n + nb - 1) / nb + 7) / 8 * 8;
  nb = (n + rows_per - 1) / rows_per;
  c.d_scratch.reserve(nb * R * P * sizeof(double));
  {
    ScopedTimer t(c, name);
    atb_partial<<<dim3((unsigned)nb, (unsigned)ctiles), 256, 0, c.stream>>>(
        A, R, B, P, n, rows_per, c.d_scratch.as<double>());
    SP_CHECK_LAUNCH();
    if ((u64)R * P <= 1024)
      sum_partials_wave<<<(unsigned)(((u64)R * P + 3) / 4), 256, 0, c.stream>>>(
          c.d_scratch.as<double>(), nb, (u64)R * P, out);
    else if ((u64)R * P <= 65536)
      sum_partials_cols<<<(unsigned)(((u64)R * P + 63) / 64), 256, 0, c.stream>>>(
          c.d_scratch.as<double>(), nb, (u64)R * P, out);
    else
      sum_partials<<<grid_for((u64)R * P, 256), 256, 0, c.stream>>>(
          c.d_scratch.as<double>(), nb, (u64)R * P, out);
    SP_CHECK_LAUNCH();
  }
}

// The R x R step of orthonormalize_gram's symmetric pass on the device
// (R <= kOrthMaxR, one wavefront), so that it needs no host round trip:
// cyclic Jacobi on (G + G^T)/2 with the round-robin ordering, R'/2 disjoint
// rotations per round (R' = R rounded up to even; an odd R's extra index
// takes none) and the host solver's stopping rule (off-diagonal sum of
// squares at most 1e-30 of the total, at most 100 sweeps); then the
// eigenvalues in descending order, the count of those above kGramRankTol of
// the largest, and M = sum over those of v v^T / sqrt(w) written over G.  On
// the nearly orthonormal factor this pass sees, G is within rounding of the
// identity and one or two sweeps suffice.  A round writes A' = J^T A J and
// V' = V J element by element into the other of two buffers: with
// row i' = alpha_i row i + beta_i row pi_i (pi_i: i's partner),
// A'[i][j] = alpha_i (alpha_j A[i][j] + beta_j A[i][pi_j])
//          + beta_i (alpha_j A[pi_i][j] + beta_j A[pi_i][pi_j]).
constexpr int kOrthMaxR = 16;

__global__ void __launch_bounds__(64) orth_small_sym(double *__restrict__ g, int R) {
  constexpr int S = kOrthMaxR + 1;  // row stride
  __shared__ double A[2][kOrthMaxR * S], V[2][kOrthMaxR * S];
  __shared__ double al[kOrthMaxR], be[kOrthMaxR], w[kOrthMaxR];
  __shared__ int pi[kOrthMaxR], ord[kOrthMaxR], ngood;
  const int l = threadIdx.x;
  for (int e = l; e < kOrthMaxR * kOrthMaxR; e += 64) {
    const int i = e / kOrthMaxR, j = e % kOrthMaxR;
    A[0][i * S + j] = i < R && j < R ? 0.5 * (g[i * R + j] + g[j * R + i]) : 0.0;
    V[0][i * S + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  const int Rp = R + (R & 1), half = Rp / 2;
  int cur = 0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = l; e < R * R; e += 64) {
      const int i = e / R, j = e % R;
      const double a = A[cur][i * S + j];
      tot += a * a;
      if (i != j) off += a * a;
    }
    off = wave_sum(off);
    tot = wave_sum(tot);
    if (off <= 1e-30 * tot || off == 0.0) break;  // the same in every lane
    for (int rnd = 0; rnd < Rp - 1; ++rnd) {
      const double *Ac = A[cur], *Vc = V[cur];
      if (l < half) {
        // pairs of the round: (rnd, R'-1) and (rnd + l, rnd - l) mod R'-1
        int p = l == 0 ? rnd : (rnd + l) % (Rp - 1);
        int q = l == 0 ? Rp - 1 : (rnd - l + Rp - 1) % (Rp - 1);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        if (q < R && Ac[p * S + q] != 0.0) {
          // t = tan of the angle that zeroes A[p][q]: the host's
          // sgn(theta) / (|theta| + sqrt(theta^2 + 1)), theta = tau / sigma,
          // with |sigma| multiplied into numerator and denominator
          const double tau = Ac[q * S + q] - Ac[p * S + p], sigma = 2.0 * Ac[p * S + q];
          const double t = (tau >= 0.0 ? sigma : -sigma) / (fabs(tau) + sqrt(tau * tau + sigma * sigma));
          c = 1.0 / sqrt(t * t + 1.0);
          sn = t * c;
        }
        al[p] = c;
        be[p] = -sn;
        pi[p] = q < R ? q : p;
        if (q < R) {
          al[q] = c;
          be[q] = sn;
          pi[q] = p;
        }
      }
      __syncthreads();
      double *An = A[cur ^ 1], *Vn = V[cur ^ 1];
      for (int e = l; e < R * R; e += 64) {
        const int i = e / R, j = e % R;
        const int pi_i = pi[i], pi_j = pi[j];
        const double ai = al[i], bi = be[i], aj = al[j], bj = be[j];
        const double x = aj * Ac[i * S + j] + bj * Ac[i * S + pi_j];
        const double y = aj * Ac[pi_i * S + j] + bj * Ac[pi_i * S + pi_j];
        An[i * S + j] = ai * x + bi * y;
        Vn[i * S + j] = aj * Vc[i * S + j] + bj * Vc[i * S + pi_j];
      }
      cur ^= 1;
      __syncthreads();
    }
  }
  const double *Ac = A[cur], *Vc = V[cur];
  // descending order; a strict total order (NaN lowest, ties by index), so
  // the ranks are a permutation
  if (l < R) {
    const double a = Ac[l * S + l], ka = a == a ? a : -HUGE_VAL;
    int rank = 0;
    for (int j = 0; j < R; ++j) {
      const double b = Ac[j * S + j], kb = b == b ? b : -HUGE_VAL;
      rank += kb > ka || (kb == ka && j < l);
    }
    ord[rank] = l;
    w[rank] = a;
  }
  __syncthreads();
  if (l == 0) {
    int n = 0;
    while (n < R && isfinite(w[n]) && w[n] > 0.0 && w[n] > w[0] * kGramRankTol) ++n;
    ngood = n;
  }
  __syncthreads();
  const int ng = ngood;
  for (int e = l; e < R * R; e += 64) {  // M[r][s], r = output column
    const int r = e / R, s = e % R;
    double v = 0.0;
    for (int q = 0; q < ng; ++q) v += Vc[s * S + ord[q]] * Vc[r * S + ord[q]] / sqrt(w[q]);
    g[e] = v;
  }
}

// U <- U V Lambda^{-1/2} from the R x R Gram U^T U: the left singular
// vectors of U in descending order (first pass), or U <- U (U^T U)^{-1/2}
// (symmetric pass, which keeps the column order of a nearly orthonormal U).
// Orthonormalises the columns of U through the eigen-decomposition of its
// R x R Gram (U <- U V w^{-1/2}, or the symmetric U V w^{-1/2} V^T).
// Directions whose Gram eigenvalue is below kGramRankTol of the largest are
// dropped and come out as zero columns; returns the number kept.
u32 orthonormalize_gram(Ctx &c, DevBuf &tmp, bool symmetric) {
  const u32 R = c.R;
  const u64 n = c.dim;
  DevBuf &g = c.ws_rr;
  g.reserve((u64)R * R * sizeof(double));
  atb(c, c.d_U.as<double>(), R, c.d_U.as<double>(), R, n, g.as<double>(), "gram_u");
  std::vector<double> G((size_t)R * R);
  SP_HIP(hipMemcpyAsync(G.data(), g.p, G.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> w, V;
  jacobi_eig(G, (int)R, w, V);
  u32 good = 0;
  while (good < R && std::isfinite(w[good]) && w[good] > 0.0 && w[good] > w[0] * kGramRankTol)
    ++good;
  // M[r][s] (R x R, "r" = output column) = V[s][r] / sqrt(w[r])
  std::vector<double> M((size_t)R * R, 0.0);
  for (u32 r = 0; r < R; ++r)
    for (u32 s = 0; s < R; ++s) {
      if (!symmetric) {
        if (r < good) M[(size_t)r * R + s] = V[(size_t)s * R + r] / std::sqrt(w[r]);
      } else {
        double v = 0.0;
        for (u32 q = 0; q < good; ++q)
          v += V[(size_t)s * R + q] * V[(size_t)r * R + q] / std::sqrt(w[q]);
        M[(size_t)r * R + s] = v;
      }
    }
  c.h_small.swap(M);
  SP_HIP(hipMemcpyAsync(g.p, c.h_small.data(), c.h_small.size() * sizeof(double),
                        hipMemcpyHostToDevice, c.stream));
  tmp.reserve(n * R * sizeof(double));
  y_times_mt<<<ymt_grid(n, R), 256, 0, c.stream>>>(c.d_U.as<double>(), n, R,
                                                   g.as<double>(), R, tmp.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(c.d_U.p, tmp.p, n * R * sizeof(double), hipMemcpyDeviceToDevice,
                        c.stream));
  return good;
}

// orthonormalize_gram's symmetric pass with its R x R step on the device
// (orth_small_sym, R <= kOrthMaxR): nothing waits.
void orthonormalize_sym_dev(Ctx &c, DevBuf &tmp) {
  const u32 R = c.R;
  const u64 n = c.dim;
  DevBuf &g = c.ws_rr;
  g.reserve((u64)R * R * sizeof(double));
  atb(c, c.d_U.as<double>(), R, c.d_U.as<double>(), R, n, g.as<double>(), "gram_u");
  orth_small_sym<<<1, 64, 0, c.stream>>>(g.as<double>(), (int)R);
  SP_CHECK_LAUNCH();
  tmp.reserve(n * R * sizeof(double));
  y_times_mt<<<ymt_grid(n, R), 256, 0, c.stream>>>(c.d_U.as<double>(), n, R,
                                                   g.as<double>(), R, tmp.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(c.d_U.p, tmp.p, n * R * sizeof(double), hipMemcpyDeviceToDevice,
                        c.stream));
}

// Host: L^{-1} (row-major, lower) of the Cholesky factor of the R x R Gram G
// (row-major).  False when G is not safely positive definite (a pivot below
// 1e-14 of the largest diagonal, or a non-finite diagonal).
bool chol_inverse(const std::vector<double> &G, u32 R, std::vector<double> &Li) {
  double dmax = 0.0;
  for (u32 i = 0; i < R; ++i) dmax = std::max(dmax, G[(size_t)i * R + i]);
  if (!(dmax > 0.0) || !std::isfinite(dmax)) return false;
  // lower Cholesky factor L (row-major)
  std::vector<double> L((size_t)R * R, 0.0);
  for (u32 j = 0; j < R; ++j) {
    double d = G[(size_t)j * R + j];
    for (u32 k = 0; k < j; ++k) d -= L[(size_t)j * R + k] * L[(size_t)j * R + k];
    if (!(d > dmax * 1e-14)) return false;
    const double ljj = std::sqrt(d);
    L[(size_t)j * R + j] = ljj;
    for (u32 i = j + 1; i < R; ++i) {
      double s = G[(size_t)i * R + j];
      for (u32 k = 0; k < j; ++k) s -= L[(size_t)i * R + k] * L[(size_t)j * R + k];
      L[(size_t)i * R + j] = s / ljj;
    }
  }
  // L^{-1} by forward substitution on the identity
  Li.assign((size_t)R * R, 0.0);
  for (u32 col = 0; col < R; ++col)
    for (u32 i = col; i < R; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (u32 k = col; k < i; ++k) s -= L[(size_t)i * R + k] * Li[(size_t)k * R + col];
      Li[(size_t)i * R + col] = s / L[(size_t)i * R + i];
    }
  return true;
}

// One Cholesky-QR pass, U <- U L^{-T} with U^T U = L L^T (device Gram, host
// Cholesky of the R x R Gram).  Two passes give Q of the thin QR with a
// positive diagonal in R, orthonormal to working precision while cond(U) is
// below ~1e7.  Returns false, leaving U untouched, when the Gram is not
// safely positive definite (a pivot below 1e-14 of the largest diagonal).
bool cholqr_pass(Ctx &c) {
  const u32 R = c.R;
  const u64 n = c.dim;
  DevBuf &g = c.ws_rr;
  g.reserve((u64)R * R * sizeof(double));
  atb(c, c.d_U.as<double>(), R, c.d_U.as<double>(), R, n, g.as<double>(), "gram_u");
  std::vector<double> G((size_t)R * R);
  SP_HIP(hipMemcpyAsync(G.data(), g.p, G.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> Li;
  if (!chol_inverse(G, R, Li)) return false;
  // Q = U L^{-T}: Q[i][r] = sum_s U[i][s] Linv[r][s], i.e. M = Linv for y_times_mt
  SP_HIP(hipMemcpyAsync(g.p, Li.data(), Li.size() * sizeof(double), hipMemcpyHostToDevice,
                        c.stream));
  DevBuf &tmp = c.ws_nr;
  tmp.reserve(n * R * sizeof(double));
  y_times_mt<<<ymt_grid(n, R), 256, 0, c.stream>>>(c.d_U.as<double>(), n, R,
                                                   g.as<double>(), R, tmp.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(c.d_U.p, tmp.p, n * R * sizeof(double), hipMemcpyDeviceToDevice,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  return true;
}

}  // namespace

// U's first `good` columns are orthonormal; replace the rest by an
// orthonormal completion (fixed pseudo-random columns, then thin QR, which
// keeps the span of the leading columns).  LAPACK's SVD likewise returns
// some orthonormal basis for the directions with zero singular value.
void complete_basis(Ctx &c, u32 good) {
  if (good >= c.R) return;
  const u64 n = c.dim;
  if (c.R > n) throw Error(SP_ERR_ARG, "rank exceeds dimension");
  fill_columns<<<grid_for(n * (c.R - good), 256), 256, 0, c.stream>>>(c.d_U.as<double>(), n,
                                                                     c.R, good);
  SP_CHECK_LAUNCH();
  dense_qr(c);
}

rocblas_handle blas_handle(Ctx &c) { return handle_for(c); }

void dense_atb(Ctx &c, const double *A, u32 R, const double *B, u64 P, u64 n, double *out) {
  atb(c, A, R, B, P, n, out, "atb");
}

void dense_ymt(Ctx &c, const double *Y, u64 n, u64 P, const double *M, u32 R, double *out) {
  y_times_mt<<<ymt_grid(n, R), 256, 0, c.stream>>>(Y, n, P, M, R, out);
  SP_CHECK_LAUNCH();
}

void dense_small_eig(const std::vector<double> &A, int n, std::vector<double> &w,
                     std::vector<double> &V) {
  jacobi_eig(A, n, w, V);
}

void dense_gemmtn(Ctx &c) {
  c.d_core.reserve((u64)c.R * c.P * sizeof(double));
  atb(c, c.d_U.as<double>(), c.R, c.d_Y.as<double>(), c.P, c.dim,
      c.d_core.as<double>(), "gemmtn");
  c.core_R = c.R;
}

void dense_symgemm(Ctx &c) {
  const u64 n = c.dim;
  const u32 R = c.R;
  DevBuf &M = c.ws_rp;
  DevBuf &out = c.ws_nr;
  M.reserve((u64)R * c.P * sizeof(double));
  out.reserve(n * R * sizeof(double));
  ScopedTimer t(c, "symgemm");
  scale_cols<<<grid_for((u64)R * c.P, 256), 256, 0, c.stream>>>(
      c.d_core.as<double>(), c.d_W.as<double>(), R, c.P, M.as<double>());
  SP_CHECK_LAUNCH();
  y_times_mt<<<ymt_grid(n, R), 256, 0, c.stream>>>(c.d_Y.as<double>(), n, c.P,
                                                   M.as<double>(), R, out.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(c.d_U.p, out.p, n * R * sizeof(double), hipMemcpyDeviceToDevice,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
}

void dense_gram(Ctx &c) {
  const u64 P = c.P, n = c.dim;
  const u64 nt = (P + GT - 1) / GT;
  std::vector<int2> tiles;
  for (u64 a = 0; a < nt; ++a)
    for (u64 b = a; b < nt; ++b) tiles.push_back(make_int2((int)a, (int)b));
  const u64 ntiles = tiles.size();
  // at most 2048 blocks (two rounds at four per CU), at least 128 rows
  // each, rows a multiple of GK
  u64 nk = std::max<u64>(1, std::min<u64>((n + 127) / 128, 2048 / ntiles));
  const u64 rows_per = ((n + nk - 1) / nk + GK - 1) / GK * GK;
  nk = (n + rows_per - 1) / rows_per;
  DevBuf &dt = c.ws_tiles;
  if (c.gram_tiles_P != P || dt.bytes < ntiles * sizeof(int2)) {
    dt.reserve(ntiles * sizeof(int2));
    SP_HIP(hipMemcpyAsync(dt.p, tiles.data(), ntiles * sizeof(int2), hipMemcpyHostToDevice,
                          c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));  // tiles is local
    c.gram_tiles_P = P;
  }
  c.d_scratch.reserve(nk * ntiles * GT * GT * sizeof(double));
  c.d_G.reserve(P * P * sizeof(double));
  {
    ScopedTimer t(c, "gram");
    // 8 ranges of ceil(ntiles nk / 8) items: see gram_partial's XCD mapping
    gram_partial<<<(unsigned)(8 * ((ntiles * nk + 7) / 8)), 256, 0, c.stream>>>(
        c.d_Y.as<double>(), n, P, rows_per, (u32)ntiles, (u32)nk, dt.as<int2>(),
        c.d_scratch.as<double>());
    SP_CHECK_LAUNCH();
    gram_finish<<<grid_for(ntiles * GT * GT, 256), 256, 0, c.stream>>>(
        c.d_scratch.as<double>(), nk, ntiles, dt.as<int2>(), P, c.d_W.as<double>(),
        c.d_G.as<double>());
    SP_CHECK_LAUNCH();
  }
}

// All eigenpairs of the P x P symmetric G by rocSOLVER (syevd, or syevj when
// selected): eigenvectors overwrite G column-major in ascending order,
// values in D.
void dense_eig_full(Ctx &c, double *G, u64 P, double *D, double *E, double *resid,
                    rocblas_int *info);

// Top-R eigenpairs through the tridiagonal form: sytrd_lower (or rocSOLVER
// sytrd) reduces G to T = Q^T G Q, the host finds T's R largest eigenpairs (bisection and
// inverse iteration, tridiag.cpp), and rocSOLVER ormtr maps only those R
// vectors back through Q.  The full divide-and-conquer over all P pairs and
// the P x P back-transformation of syevd are not needed.
void eig_top_tridiag(Ctx &c, double *G, u64 P, u32 R, double *V, std::vector<double> &theta) {
  rocblas_handle h = handle_for(c);
  // workspace: D (P), E (P), tau (P), R vectors (R x P)
  c.ws_eig.reserve((3 * P + (u64)R * P) * sizeof(double));
  double *D = c.ws_eig.as<double>();
  double *E = D + P;
  double *tau = E + P;
  double *Zd = tau + P;
  // the reflectors: in Vr (lower form, G only read) from the own reduction,
  // in G (upper form) from rocSOLVER's; Vr only where the own one can run
  double *Vr = nullptr;
  if (c.eig_method == 2 && P <= kTrdMaxP) {
    c.ws_trd_v.reserve(P * P * sizeof(double));
    Vr = c.ws_trd_v.as<double>();
  }
  bool lower;
  std::vector<double> &hd = c.h_eig_de, &Z = c.h_eig_z;
  hd.resize(2 * P);
  Z.resize((u64)R * P);
  {
    ScopedTimer t(c, "sytrd");
    const int *abort_flag = nullptr;
    lower = Vr && sytrd_lower(c, G, P, D, E, tau, Vr, &abort_flag);
    if (lower) {
      // the reduction's output and its time-out flag in one wait
      SP_HIP(hipMemcpyAsync(hd.data(), D, 2 * P * sizeof(double), hipMemcpyDeviceToHost,
                            c.stream));
      c.h_trd_abort = 0;
      if (abort_flag)
        SP_HIP(hipMemcpyAsync(&c.h_trd_abort, abort_flag, sizeof(int), hipMemcpyDeviceToHost,
                              c.stream));
      SP_HIP(hipStreamSynchronize(c.stream));
      if (c.h_trd_abort) {
        // some workgroup was not resident in time (the device is shared with
        // other work): every workgroup has left and G is untouched
        ++c.trd_timeouts;
        ScopedTimer mark(c, "sytrd_timeout");  // counted fallback
        lower = false;
      }
    }
    if (!lower) {
      if (rocsolver_dsytrd(h, rocblas_fill_upper, (rocblas_int)P, G, (rocblas_int)P, D, E,
                           tau) != rocblas_status_success)
        throw Error(SP_ERR_DEVICE, "rocsolver_dsytrd failed");
      SP_HIP(hipMemcpyAsync(hd.data(), D, 2 * P * sizeof(double), hipMemcpyDeviceToHost,
                            c.stream));
      SP_HIP(hipStreamSynchronize(c.stream));
    }
  }
  theta.resize(R);
  {
    ScopedTimer t(c, "tridiag_top");
    tridiag_top(hd.data(), hd.data() + P, (int)P, (int)R, theta.data(), Z.data());
  }
  SP_HIP(hipMemcpyAsync(Zd, Z.data(), Z.size() * sizeof(double), hipMemcpyHostToDevice,
                        c.stream));
  {
    ScopedTimer t(c, "ormtr");
    if (!(lower && apply_q_lower(c, Vr, tau, P, Zd, R, V))) {
      if (rocsolver_dormtr(h, rocblas_side_left, lower ? rocblas_fill_lower : rocblas_fill_upper,
                           rocblas_operation_none, (rocblas_int)P, (rocblas_int)R,
                           lower ? Vr : G, (rocblas_int)P, tau, Zd,
                           (rocblas_int)P) != rocblas_status_success)
        throw Error(SP_ERR_DEVICE, "rocsolver_dormtr failed");
      transpose_nr<<<grid_for(P * R, 256), 256, 0, c.stream>>>(Zd, R, P, V);
      SP_CHECK_LAUNCH();
    }
  }
  // V is queued on the stream; Z stays in the context's staging buffer
}

int scale_to_unit(Ctx &c, double *A, u64 count) {
  constexpr unsigned kBlocks = 256;
  c.ws_amax.reserve(kBlocks * sizeof(double));
  absmax_partial<<<kBlocks, 256, 0, c.stream>>>(A, count, c.ws_amax.as<double>());
  SP_CHECK_LAUNCH();
  double part[kBlocks];
  SP_HIP(hipMemcpyAsync(part, c.ws_amax.p, sizeof(part), hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  double amax = 0.0;
  for (double q : part) amax = std::max(amax, q);
  if (!(amax > 0.0 && std::isfinite(amax) && (amax < 0x1p-400 || amax > 0x1p400))) return 0;
  const int ex = -std::ilogb(amax);
  scale_by<<<grid_for(count, 256), 256, 0, c.stream>>>(A, count, std::ldexp(1.0, ex));
  SP_CHECK_LAUNCH();
  return ex;
}

void scale_to_unit_async(Ctx &c, double *A, u64 count) {
  constexpr unsigned kBlocks = 256;
  c.ws_amax.reserve(kBlocks * sizeof(double) + sizeof(int));
  double *part = c.ws_amax.as<double>();
  int *ex = reinterpret_cast<int *>(part + kBlocks);
  absmax_partial<<<kBlocks, 256, 0, c.stream>>>(A, count, part);
  SP_CHECK_LAUNCH();
  scale_pick<<<1, 256, 0, c.stream>>>(part, (int)kBlocks, ex);
  SP_CHECK_LAUNCH();
  scale_by_pow2<<<grid_for(count, 256), 256, 0, c.stream>>>(A, count, ex);
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(c.h_pin.get(), ex, sizeof(int), hipMemcpyDeviceToHost, c.stream));
}

void eig_top(Ctx &c, double *G, u64 P, u32 R, double *V, std::vector<double> &theta,
             bool in_range) {
  if (R > P) throw Error(SP_ERR_ARG, "rank exceeds the matrix order");
  // A Gram far outside [2^-400, 2^400] (e.g. from a nearly vanishing Y) is
  // scaled by a power of two first (exact), as LAPACK's dsyevd scales its
  // input (rocSOLVER's does not), so that Householder sums of squares and
  // pivots stay in range; the eigenvalues are scaled back.  in_range: the
  // caller has done so already.
  const int ex = in_range ? 0 : scale_to_unit(c, G, P * P);
  if (c.eig_method >= 2) {
    eig_top_tridiag(c, G, P, R, V, theta);
    for (u32 r = 0; r < R; ++r) theta[r] = std::ldexp(theta[r], -ex);
    return;
  }
  // workspace: D (P), E (P), residual (2), then int words
  c.ws_eig.reserve((2 * P + 2) * sizeof(double) + 4 * sizeof(rocblas_int));
  double *D = c.ws_eig.as<double>();
  double *E = D + P;
  double *resid = E + P;
  rocblas_int *info = reinterpret_cast<rocblas_int *>(resid + 2);
  dense_eig_full(c, G, P, D, E, resid, info);
  std::vector<double> lam(P);
  SP_HIP(hipMemcpyAsync(lam.data(), D, P * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  theta.resize(R);
  for (u32 r = 0; r < R; ++r) theta[r] = std::ldexp(lam[P - 1 - r], -ex);
  top_columns<<<grid_for(P * R, 256), 256, 0, c.stream>>>(G, P, R, V);
  SP_CHECK_LAUNCH();
}

// Generalised Rayleigh-Ritz on the span of the k rows of Kt (k x P, device),
// GKt = Kt G: H = Kt G Kt^T and M = Kt Kt^T on the device, H s = theta M s
// on the host (Cholesky of M, Jacobi), then the top R Ritz vectors as the
// rows of Yt, G Yt in GYt and their squared residuals ||G y - theta y||^2 in
// res (host).  small: 3 k^2 doubles, th_d: R, res_d: R (device).  Returns
// the largest residual / theta_1, or infinity when M is not safely positive
// definite or theta_R is not above 1e-8 theta_1.
double rayleigh_ritz(Ctx &c, const double *Kt, const double *GKt, u32 k, u64 P, u32 R,
                     double *small, double *th_d, double *res_d, double *Yt, double *GYt,
                     std::vector<double> &theta, std::vector<double> &res) {
  constexpr double kFail = std::numeric_limits<double>::infinity();
  const u64 kk = (u64)k * k;
  y_times_mt<<<ymt_grid(k, k), 256, 0, c.stream>>>(Kt, k, P, GKt, k, small);     // H
  SP_CHECK_LAUNCH();
  y_times_mt<<<ymt_grid(k, k), 256, 0, c.stream>>>(Kt, k, P, Kt, k, small + kk); // M
  SP_CHECK_LAUNCH();
  std::vector<double> HM(2 * kk);
  SP_HIP(hipMemcpyAsync(HM.data(), small, HM.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  auto sym = [k](double *A) {
    for (u32 i = 0; i < k; ++i)
      for (u32 j = 0; j < i; ++j)
        A[(size_t)i * k + j] = A[(size_t)j * k + i] =
            0.5 * (A[(size_t)i * k + j] + A[(size_t)j * k + i]);
  };
  sym(HM.data());
  sym(HM.data() + kk);
  std::vector<double> M(HM.begin() + kk, HM.end()), Li;
  if (!chol_inverse(M, k, Li)) return kFail;
  // C = L^{-1} H L^{-T} (symmetrised), C = S diag(w) S^T
  std::vector<double> T1(kk, 0.0), C(kk, 0.0);
  for (u32 i = 0; i < k; ++i)
    for (u32 j = 0; j < k; ++j) {
      double v = 0.0;
      for (u32 q = 0; q <= i; ++q) v += Li[(size_t)i * k + q] * HM[(size_t)q * k + j];
      T1[(size_t)i * k + j] = v;
    }
  for (u32 i = 0; i < k; ++i)
    for (u32 j = 0; j < k; ++j) {
      double v = 0.0;
      for (u32 q = 0; q <= j; ++q) v += T1[(size_t)i * k + q] * Li[(size_t)j * k + q];
      C[(size_t)i * k + j] = v;
    }
  sym(C.data());
  std::vector<double> w, S;
  jacobi_eig(C, (int)k, w, S);
  if (!(w[0] > 0.0) || !std::isfinite(w[0]) || !(w[R - 1] > w[0] * 1e-8)) return kFail;
  // Ritz vectors y_r = Kt^T b_r with B = L^{-T} S (first R columns): Yt = B^T Kt
  std::vector<double> B((size_t)k * R, 0.0);
  for (u32 s = 0; s < k; ++s)
    for (u32 r = 0; r < R; ++r) {
      double v = 0.0;
      for (u32 q = s; q < k; ++q) v += Li[(size_t)q * k + s] * S[(size_t)q * k + r];
      B[(size_t)s * R + r] = v;
    }
  double *Bd = small + 2 * kk;
  SP_HIP(hipMemcpyAsync(Bd, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice,
                        c.stream));
  SP_HIP(hipMemcpyAsync(th_d, w.data(), R * sizeof(double), hipMemcpyHostToDevice, c.stream));
  atb(c, Bd, R, Kt, P, k, Yt, "warm_atb");
  atb(c, Bd, R, GKt, P, k, GYt, "warm_atb");
  resid_sq<<<R, 256, 0, c.stream>>>(GYt, Yt, th_d, P, res_d);
  SP_CHECK_LAUNCH();
  res.resize(R);
  SP_HIP(hipMemcpyAsync(res.data(), res_d, R * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));  // also the point after which B and w may go
  double worst = 0.0;
  for (u32 r = 0; r < R; ++r) worst = std::max(worst, std::sqrt(res[r]));
  theta.assign(w.begin(), w.begin() + R);
  return worst / w[0];
}

// Top-R eigenpairs of the Gram G (P x P, device, in range after
// scale_to_unit) by Rayleigh-Ritz on span(X0), X0 = Z^T U =
// diag(sqrt W) (U^T Y)^T with U the factor Y was computed from, and, when
// that misses by less than 1e-3, on span(X0, R0) with R0 the first step's
// Ritz residuals G y - theta y (orthogonal to span(X0); one block Krylov
// step).  In the Tucker loops U is the previous iteration's answer: X0's
// columns are G's top-R eigenvectors times the singular values once the
// iteration has reached its fixed point, and converge to them at the loop's
// own rate before (one HOOI step is one subspace iteration from U).  Returns
// the largest Ritz residual ||G y - theta y|| / theta_1 of the last step
// (infinity when a basis is rank-deficient or theta_R is near the rank
// floor); V (P x R row-major) and theta (descending) hold its Ritz pairs,
// which the caller keeps only when that residual is at the accuracy of the
// direct solver.
double eig_top_warm(Ctx &c, const double *G, u64 P, u32 R, double *V,
                    std::vector<double> &theta) {
  const u64 n = c.dim, RP = (u64)R * P, RR = (u64)R * R;
  // Xt, GXt, Yt, GYt: [R][P]; Kt, GKt: [2R][P]; Y2, GY2: [R][P]; small:
  // 3 (2R)^2; theta, residuals, inverse residual norms: R each
  c.ws_warm.reserve((10 * RP + 12 * RR + 3 * R) * sizeof(double));
  double *Xt = c.ws_warm.as<double>();
  double *GXt = Xt + RP;
  double *Yt = GXt + RP;
  double *GYt = Yt + RP;
  double *Kt = GYt + RP;
  double *GKt = Kt + 2 * RP;
  double *Y2 = GKt + 2 * RP;
  double *GY2 = Y2 + RP;
  double *small = GY2 + RP;
  double *th_d = small + 12 * RR;
  double *res_d = th_d + R;
  double *inv_d = res_d + R;
  atb(c, c.d_U.as<double>(), R, c.d_Y.as<double>(), P, n, Yt, "warm_x0");  // U^T Y
  scale_cols_sqrt<<<grid_for(RP, 256), 256, 0, c.stream>>>(Yt, c.d_W.as<double>(), R, P, Xt);
  SP_CHECK_LAUNCH();
  y_times_mt<<<ymt_grid(R, P), 256, 0, c.stream>>>(Xt, R, P, G, P, GXt);  // rows of X0^T G
  SP_CHECK_LAUNCH();
  std::vector<double> res;
  double worst = rayleigh_ritz(c, Xt, GXt, R, P, R, small, th_d, res_d, Yt, GYt, theta, res);
  const double *Yfin = Yt;
  if (worst > 1e-13 && worst < 1e-3 && 2 * (u64)R <= P) {
    // second block: the Ritz residual directions, each scaled to unit length
    std::vector<double> inv(R);
    for (u32 r = 0; r < R; ++r) inv[r] = 1.0 / std::max(std::sqrt(res[r]), 1e-300);
    SP_HIP(hipMemcpyAsync(inv_d, inv.data(), R * sizeof(double), hipMemcpyHostToDevice,
                          c.stream));
    SP_HIP(hipMemcpyAsync(Kt, Yt, RP * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    SP_HIP(hipMemcpyAsync(GKt, GYt, RP * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    ritz_dir<<<grid_for(RP, 256), 256, 0, c.stream>>>(GYt, Yt, th_d, inv_d, R, P, Kt + RP);
    SP_CHECK_LAUNCH();
    y_times_mt<<<ymt_grid(R, P), 256, 0, c.stream>>>(Kt + RP, R, P, G, P, GKt + RP);
    SP_CHECK_LAUNCH();
    worst = rayleigh_ritz(c, Kt, GKt, 2 * R, P, R, small, th_d, res_d, Y2, GY2, theta, res);
    Yfin = Y2;  // inv is read by the upload above; rayleigh_ritz synchronised
  }
  transpose_nr<<<grid_for(RP, 256), 256, 0, c.stream>>>(Yfin, R, P, V);
  SP_CHECK_LAUNCH();
  return worst;
}

// Top-R eigenpairs of the Gram G (P <= 256, in range after scale_to_unit) by
// Rayleigh-Ritz on the m-block Krylov space of X0 = Z^T U (the warm start of
// eig_top_warm): krylov_basis builds orthonormal rows Q and Q G in one
// workgroup, H = Q G Q^T (m R x m R) goes through the direct solver (eig_top),
// and the Ritz vectors Y = S^T Q come with their residuals ||G y - theta y||.
// Returns the largest residual / theta_1 (infinity when theta_R is not above
// 1e-8 theta_1 or the basis could not be built); V and theta as
// eig_top_warm's.
double eig_top_krylov(Ctx &c, const double *G, u64 P, u32 R, u32 m, double *V,
                      std::vector<double> &theta) {
  constexpr double kFail = std::numeric_limits<double>::infinity();
  const u64 n = c.dim, RP = (u64)R * P, k = (u64)m * R;
  // Xt, Yt, GYt: [R][P]; Qt, GQt: [k][P]; H: [k][k]; S: [k][R]; theta, residuals: R
  c.ws_warm.reserve((3 * RP + 2 * k * P + k * k + k * R + 2 * R) * sizeof(double));
  double *Xt = c.ws_warm.as<double>();
  double *Yt = Xt + RP;
  double *GYt = Yt + RP;
  double *Qt = GYt + RP;
  double *GQt = Qt + k * P;
  double *H = GQt + k * P;
  double *S = H + k * k;
  double *th_d = S + k * R;
  double *res_d = th_d + R;
  atb(c, c.d_U.as<double>(), R, c.d_Y.as<double>(), P, n, Yt, "warm_x0");  // U^T Y
  scale_cols_sqrt<<<grid_for(RP, 256), 256, 0, c.stream>>>(Yt, c.d_W.as<double>(), R, P, Xt);
  SP_CHECK_LAUNCH();
  if (!krylov_basis(c, G, Xt, P, R, m, Qt, GQt)) return kFail;
  y_times_mt<<<ymt_grid(k, k), 256, 0, c.stream>>>(Qt, k, P, GQt, k, H);
  SP_CHECK_LAUNCH();
  std::vector<double> th;
  eig_top(c, H, k, R, S, th, true);
  if (!(th[0] > 0.0) || !std::isfinite(th[0]) || !(th[R - 1] > th[0] * 1e-8)) return kFail;
  atb(c, S, R, Qt, P, k, Yt, "warm_atb");
  atb(c, S, R, GQt, P, k, GYt, "warm_atb");
  c.h_small.assign(th.begin(), th.end());
  SP_HIP(hipMemcpyAsync(th_d, c.h_small.data(), R * sizeof(double), hipMemcpyHostToDevice,
                        c.stream));
  resid_sq<<<R, 256, 0, c.stream>>>(GYt, Yt, th_d, P, res_d);
  SP_CHECK_LAUNCH();
  std::vector<double> res(R);
  SP_HIP(hipMemcpyAsync(res.data(), res_d, R * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  double worst = 0.0;
  for (u32 r = 0; r < R; ++r) worst = std::max(worst, std::sqrt(res[r]));
  transpose_nr<<<grid_for(RP, 256), 256, 0, c.stream>>>(Yt, R, P, V);
  SP_CHECK_LAUNCH();
  theta = th;
  return worst / th[0];
}

void dense_leftsvd_y(Ctx &c, double *sigma) {
  dense_gram(c);
  const u64 P = c.P, n = c.dim;
  const u32 R = c.R;
  if (R > P) throw Error(SP_ERR_ARG, "rank exceeds packed width");
  // top-R eigenvectors of G as V (P x R, row-major), eigenvalues descending:
  // the warm-started Rayleigh-Ritz step when it converges, else the direct
  // solver (both on G scaled into range; theta scaled back)
  DevBuf &V = c.ws_pr;
  V.reserve(P * R * sizeof(double));
  std::vector<double> theta;
  // G scaled into range on the device; the exponent is read after the
  // eigen-step below, every path of which synchronises the stream
  scale_to_unit_async(c, c.d_G.as<double>(), P * P);
  bool warm = false;
  // the forms are timed between events on the stream: the time from when the
  // stream reaches the step (after the queued S3TTMc and Gram) to its end
  hipEvent_t *ev = c.eig_ev.get();
  auto ms_between = [](hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    SP_HIP(hipEventElapsedTime(&ms, a, b));
    return (double)ms;
  };
  auto mean = [](double m, double x) { return m > 0.0 ? 0.75 * m + 0.25 * x : x; };
  if (c.warm_enabled && c.eig_method == 2) {
    if (c.warm_skip > 0) {
      --c.warm_skip;
    } else {
      double res;
      SP_HIP(hipEventRecord(ev[0], c.stream));
      // P <= 256: the block Krylov form; above, the one- or two-block step.
      // Its number of blocks m adapts: two more after a miss, and that m+1 is
      // then the floor for the next 64 calls; one fewer after a residual two
      // digits inside the bar while above the floor
      const u32 mmax = (u32)std::min<u64>(12, P / R);
      const bool kry = P <= 256 && R <= 16 && mmax >= 2;
      if (kry && c.kry_m == 0) c.kry_m = std::min<u32>(8, mmax);
      if (kry && c.kry_m > mmax) c.kry_m = mmax;
      const bool widest = kry && c.kry_m >= mmax;
      {
        ScopedTimer t(c, "eig_warm");
        res = kry ? eig_top_krylov(c, c.d_G.as<double>(), P, R, c.kry_m, V.as<double>(), theta)
                  : eig_top_warm(c, c.d_G.as<double>(), P, R, V.as<double>(), theta);
      }
      SP_HIP(hipEventRecord(ev[1], c.stream));
      if (kry) {
        if (res > 1e-13) {
          c.kry_floor = std::min(mmax, c.kry_m + 1);
          c.kry_hold = 64;
          c.kry_m = std::min(mmax, c.kry_m + 2);
        } else {
          if (c.kry_hold > 0 && --c.kry_hold == 0) c.kry_floor = 2;
          if (res <= 1e-15 && c.kry_m > std::max<u32>(2, c.kry_floor)) --c.kry_m;
        }
      }
      // kept at the direct solver's accuracy.  A near miss (the loop is
      // closing in on its fixed point) is tried again at the next call;
      // after a clear miss (or a miss of the Krylov form at its most blocks,
      // where its own rounding can keep the residual just above the bar) the
      // next calls skip the attempt, as many as keep the misses within a
      // twentieth of the direct path's time
      warm = res <= 1e-13;
      if (warm) {
        ScopedTimer mark(c, "eig_warm_ok");  // counted success
      } else {
        SP_HIP(hipEventSynchronize(ev[1]));
        c.warm_miss_ms = mean(c.warm_miss_ms, ms_between(ev[0], ev[1]));
        if (res > 1e-6 || widest)
          c.warm_skip = c.direct_eig_ms > 0.0
                            ? (u32)std::min(64.0, std::ceil(c.warm_miss_ms /
                                                            (0.05 * c.direct_eig_ms)) - 1.0)
                            : 1;
      }
    }
  }
  if (!warm) {
    SP_HIP(hipEventRecord(ev[2], c.stream));
    eig_top(c, c.d_G.as<double>(), P, R, V.as<double>(), theta, true);
    SP_HIP(hipEventRecord(ev[3], c.stream));
    SP_HIP(hipEventSynchronize(ev[3]));
    c.direct_eig_ms = mean(c.direct_eig_ms, ms_between(ev[2], ev[3]));
  }
  const int ex = c.h_pin.get()[0];
  for (u32 r = 0; r < R; ++r) theta[r] = std::ldexp(theta[r], -ex);
  c.ws_eig.reserve((2 * P + 2 + R) * sizeof(double) + 4 * sizeof(rocblas_int));
  double *theta_d = c.ws_eig.as<double>();
  if (sigma)
    for (u32 r = 0; r < R; ++r) sigma[r] = std::sqrt(std::max(theta[r], 0.0));
  // The Gram's eigenvalues carry absolute errors of about eps * lambda_max,
  // so directions with lambda below kGramRankTol * lambda_max (singular value
  // below ~3e-6 of the largest), and all of them for a zero Y, have no
  // reliable left vector; they get an orthonormal completion
  u32 good = 0;
  while (good < R && theta[good] > 0.0 && theta[good] > theta[0] * kGramRankTol) ++good;
  const double lmax = std::max(std::fabs(theta[0]), 1e-300);
  c.h_theta.assign(theta.begin(), theta.end());
  SP_HIP(hipMemcpyAsync(theta_d, c.h_theta.data(), R * sizeof(double), hipMemcpyHostToDevice,
                        c.stream));
  DevBuf &M = c.ws_rp;
  M.reserve((u64)R * P * sizeof(double));
  svd_map<<<grid_for((u64)R * P, 256), 256, 0, c.stream>>>(
      V.as<double>(), theta_d, c.d_W.as<double>(), P, R, lmax * 1e-30, M.as<double>());
  SP_CHECK_LAUNCH();
  {
    ScopedTimer t(c, "y_times_v");
    y_times_mt<<<ymt_grid(n, R), 256, 0, c.stream>>>(c.d_Y.as<double>(), n, P,
                                                     M.as<double>(), R, c.d_U.as<double>());
    SP_CHECK_LAUNCH();
  }
  complete_basis(c, good);
}

void dense_eig_full(Ctx &c, double *G, u64 P, double *D, double *E, double *resid,
                    rocblas_int *info) {
  rocblas_handle h = handle_for(c);
  if (c.eig_method == 1) {
    ScopedTimer t(c, "syevj");
    rocblas_status st = rocsolver_dsyevj(h, rocblas_esort_ascending, rocblas_evect_original,
                                         rocblas_fill_upper, (rocblas_int)P,
                                         G, (rocblas_int)P, 0.0, resid, 100,
                                         info + 1, D, info);
    if (st != rocblas_status_success) throw Error(SP_ERR_DEVICE, "rocsolver_dsyevj failed");
  } else {
    ScopedTimer t(c, "syevd");
    rocblas_status st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper,
                                         (rocblas_int)P, G, (rocblas_int)P, D, E, info);
    if (st != rocblas_status_success) throw Error(SP_ERR_DEVICE, "rocsolver_dsyevd failed");
  }
  rocblas_int hinfo = 0;
  SP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  if (hinfo != 0) throw Error(SP_ERR_NUMERIC, "eigensolver did not converge");
}

// Thin QR of U (matrix::qr, matrixops.cpp:58-75).  CholeskyQR2 (two Gram +
// R x R Cholesky passes) gives the same Q as Householder up to column signs
// (the sign/rotation freedom of quirk A14) at a fraction of rocSOLVER's
// launch count; ill-conditioned U falls back to rocSOLVER geqrf + orgqr.
void dense_qr(Ctx &c) {
  ScopedTimer t(c, "qr");
  if (cholqr_pass(c) && cholqr_pass(c)) return;
  const u64 n = c.dim;
  const u32 R = c.R;
  DevBuf &At = c.ws_nr;
  DevBuf &tau = c.ws_rr;
  At.reserve(n * R * sizeof(double));
  tau.reserve((u64)R * R * sizeof(double));
  transpose_nr<<<grid_for(n * R, 256), 256, 0, c.stream>>>(c.d_U.as<double>(), n, R,
                                                           At.as<double>());
  SP_CHECK_LAUNCH();
  rocblas_handle h = handle_for(c);
  if (rocsolver_dgeqrf(h, (rocblas_int)n, (rocblas_int)R, At.as<double>(), (rocblas_int)n,
                       tau.as<double>()) != rocblas_status_success)
    throw Error(SP_ERR_DEVICE, "rocsolver_dgeqrf failed");
  if (rocsolver_dorgqr(h, (rocblas_int)n, (rocblas_int)R, (rocblas_int)R, At.as<double>(),
                       (rocblas_int)n, tau.as<double>()) != rocblas_status_success)
    throw Error(SP_ERR_DEVICE, "rocsolver_dorgqr failed");
  transpose_nr<<<grid_for(n * R, 256), 256, 0, c.stream>>>(At.as<double>(), R, n,
                                                           c.d_U.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipStreamSynchronize(c.stream));
}

void dense_regularize(Ctx &c, double delta) {
  const u64 n = c.dim;
  clip_rows<<<grid_for(n, 256), 256, 0, c.stream>>>(c.d_U.as<double>(), n, c.R, delta);
  SP_CHECK_LAUNCH();
  // first pass on the host: its count of kept directions decides the
  // completion; the second pass sees orthonormal columns
  complete_basis(c, orthonormalize_gram(c, c.ws_nr, false));
  if (c.R <= (u32)kOrthMaxR)
    orthonormalize_sym_dev(c, c.ws_nr);
  else
    orthonormalize_gram(c, c.ws_nr, true);
}

void dense_factor_snapshot(Ctx &c) {
  const u64 bytes = (u64)c.dim * c.R * sizeof(double);
  c.ws_uprev.reserve(bytes);
  SP_HIP(hipMemcpyAsync(c.ws_uprev.p, c.d_U.p, bytes, hipMemcpyDeviceToDevice, c.stream));
  c.uprev_R = c.R;
}

// sin of the largest principal angle between span(U) and span(U_snapshot),
// both with orthonormal columns: sqrt(1 - sigma_min(U^T U_snapshot)^2), with
// sigma_min^2 the smallest eigenvalue of M^T M, M = U^T U_snapshot (R x R,
// device product, host Jacobi).  Resolution about 1e-8 (the 1 - c^2 form).
double dense_subspace_sin(Ctx &c) {
  const u32 R = c.R;
  DevBuf &g = c.ws_rr;
  g.reserve((u64)R * R * sizeof(double));
  atb(c, c.d_U.as<double>(), R, c.ws_uprev.as<double>(), R, c.dim, g.as<double>(),
      "subspace_atb");
  std::vector<double> M((size_t)R * R);
  SP_HIP(hipMemcpyAsync(M.data(), g.p, M.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> MtM((size_t)R * R, 0.0);
  for (u32 i = 0; i < R; ++i)
    for (u32 j = 0; j < R; ++j) {
      double v = 0.0;
      for (u32 k = 0; k < R; ++k) v += M[(size_t)k * R + i] * M[(size_t)k * R + j];
      MtM[(size_t)i * R + j] = v;
    }
  std::vector<double> w, V;
  jacobi_eig(MtM, (int)R, w, V);
  const double smin2 = std::min(1.0, std::max(0.0, w[R - 1]));
  return std::sqrt(1.0 - smin2);
}

double dense_core_norm(Ctx &c) {
  std::vector<double> core((u64)c.R * c.P), W(c.P);
  SP_HIP(hipMemcpyAsync(core.data(), c.d_core.p, core.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipMemcpyAsync(W.data(), c.d_W.p, W.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (u32 r = 0; r < c.R; ++r)
    for (u64 p = 0; p < c.P; ++p) s += std::pow(core[(u64)r * c.P + p], 2) * W[p];
  return std::sqrt(s);
}

}  // namespace spamd
