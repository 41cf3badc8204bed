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
// Row r of A is produced by one wavefront from the owners whose rem is r
// (the rem-sorted work list, leaves ascending), scattering into an LDS tile of
// A[r, c0 : c0+TILE].  Contributions are added in ascending leaf order, each
// as mult * (v_r * v_c), so A[r][c] and A[c][r] are the same floating-point
// sum and A is exactly symmetric; the result is deterministic.
#include "sp_internal.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

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

void hoevd_init(Ctx &c, u32 R, double *U_out, double *eig_out) {
  const u64 n = c.dim;
  if (R < 1 || R > n) throw Error(SP_ERR_ARG, "rank must be in [1, dim]");
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
