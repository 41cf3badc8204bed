// Block Krylov basis for the eigen-step's top-R eigenpairs of a Gram with
// P <= 256 (the step replaces matrix::syevd in leftsvd,
// src/dense/matrixops.cpp:18-43 and 77-100, when it reaches the direct
// solver's accuracy; the caller checks that).
//
// One workgroup builds the orthonormal rows Q (k x P, k = m b) of the block
// Krylov space span{X0, G X0, ..., G^(m-1) X0} of the b starting rows X0
// (the warm start Z^T U of leftsvd) and the rows of Q G:
//   Q_0 = CholQR2(X0);  for j = 0..m-1:  (Q G)_j = Q_j G,  and for j < m-1
//   Q_{j+1} = CholQR(proj(CholQR(proj(Z)))) with Z = (Q G)_j and proj the
//   classical Gram-Schmidt projection off Q_0..Q_j: the second projection
//   follows the first normalisation, so the block is orthogonal to the
//   earlier ones to working precision even where the first normalisation
//   scaled up a direction of small norm.
// A block whose projected norm falls below 1e-10 of its norm before the
// projection means the space is invariant to working precision: the basis
// stops there and the remaining rows of Q and Q G are zero (they add zero
// eigenvalues to Q G Q^T, which the caller's top R do not reach).  A start
// block that is not of full rank leaves everything zero, and the caller's
// residual test then fails.
//
// Layout: G, Q and Q G in global memory (P <= 256 rows, L2-resident), the
// current block Z in the LDS as rows, the previous block transposed for the
// product with G.  Every sum has a fixed order, so the result is
// deterministic.
#include "sp_internal.hpp"
#include "wave.hpp"

namespace spamd {

namespace {

constexpr int kKyThreads = 512;
constexpr int kKyWaves = kKyThreads / 64;
constexpr int kKyMaxP = 256;

struct KyArgs {
  const double *G;    // P x P, symmetric
  const double *X0t;  // b x P: the starting block, as rows
  double *Qt;         // (m b) x P: orthonormal rows, block j in rows j b ..
  double *GQt;        // (m b) x P: rows of Q G
  int P, b, m;
};

__device__ inline void ky_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Sum over the wave, the same bits in every lane (pairs that add the same
// two operands: quad swaps, half-row and row mirrors, then the half-row and
// half-wave swaps).
template <int C>
__device__ inline double ky_dpp(double x) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)v, C, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), C, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double ky_swap_sum(double x, bool wide) {
  const long long v = __double_as_longlong(x);
  const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  const auto a = wide ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                      : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto b = wide ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                      : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __longlong_as_double(((long long)b[0] << 32) | a[0]) +
         __longlong_as_double(((long long)b[1] << 32) | a[1]);
}
__device__ inline double ky_wave_sum(double x) {
  x += ky_dpp<0xb1>(x);
  x += ky_dpp<0x4e>(x);
  x += ky_dpp<0x141>(x);
  x += ky_dpp<0x140>(x);
  return ky_swap_sum(ky_swap_sum(x, false), true);
}

typedef double kd4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64: lane l supplies A[m = l%16][k = l/16] and
// B[k = l/16][n = l%16]; accumulator register q holds C[m = 4q + l/16][n = l%16].
__device__ inline kd4 ky_mfma(double a, double b, kd4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Z = Q_j G (rows r < b of Zb, LDS) from QbT[i][r] = Q_j[r][i] (LDS): column
// tile ct (16 columns) on wave ct mod 8, K = the P rows of G in steps of 4,
// two accumulators.  Z also goes to GQt rows row0 ..
template <int B>
__device__ void ky_gprod(const KyArgs &a, int row0, double *Zb, const double *QbT) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int kr = lane >> 4, kc = lane & 15, P = a.P, nt = (P + 15) / 16;
  // loads are unconditional (indices clamped into range, the values then
  // selected), so the loop has no branch and its loads can run ahead; rows of
  // QbT past P are zero
  const bool kin = kc < B;
  const double *qb = QbT + (kc < B ? kc : 0);
  for (int ct = wave; ct < nt; ct += kKyWaves) {
    const int c = ct * 16 + kc;
    const bool cin = c < P;
    const double *gc = a.G + (cin ? c : P - 1);
    kd4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < P; k0 += 32) {  // past P: zeros (QbT has 256 rows)
      double ga[8], qa[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = k0 + 4 * u + kr;
        ga[u] = gc[(size_t)(kk < P ? kk : P - 1) * P];
        qa[u] = qb[(kk < kKyMaxP ? kk : kKyMaxP - 1) * B];
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        const int k1 = k0 + 4 * u + kr, k2 = k1 + 4;
        acc0 = ky_mfma(kin && k1 < P ? qa[u] : 0.0, k1 < P && cin ? ga[u] : 0.0, acc0);
        acc1 = ky_mfma(kin && k2 < P ? qa[u + 1] : 0.0, k2 < P && cin ? ga[u + 1] : 0.0, acc1);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 4 * q + kr;
      if (r < B) {
        const double z = acc0[q] + acc1[q];
        Zb[r * kKyMaxP + c] = z;
        if (r < a.b && c < P) a.GQt[(size_t)(row0 + r) * P + c] = z;
      }
    }
  }
  ky_barrier();
}

// Z -= Q_{0..k-1}^T (Q Z^T): the coefficients C = Q Z^T (k x B; row tile qt
// on wave qt mod 8, K = the P columns) into Cm, then Z -= C^T Q (column tile
// ct on wave ct mod 8, K = k); each entry of Z is owned by one lane.
template <int B>
__device__ void ky_project(const KyArgs &a, int k, double *Zb, double *Cm) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int kr = lane >> 4, kc = lane & 15, P = a.P;
  // branch-free loops as in ky_gprod; columns of Zb past P are zero
  const bool kin = kc < B;
  const double *zr = Zb + (kc < B ? kc : 0) * kKyMaxP;
  for (int qt = wave; qt * 16 < k; qt += kKyWaves) {
    const int q = qt * 16 + kc;
    const bool qin = q < k;
    const double *qrow = a.Qt + (size_t)(qin ? q : k - 1) * P;
    kd4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    for (int c0 = 0; c0 < P; c0 += 32) {
      double xa[8], za[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int cc = c0 + 4 * u + kr;
        xa[u] = qrow[cc < P ? cc : P - 1];
        za[u] = zr[cc < kKyMaxP ? cc : kKyMaxP - 1];
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        const int c1 = c0 + 4 * u + kr, c2 = c1 + 4;
        acc0 = ky_mfma(qin && c1 < P ? xa[u] : 0.0, kin && c1 < P ? za[u] : 0.0, acc0);
        acc1 = ky_mfma(qin && c2 < P ? xa[u + 1] : 0.0, kin && c2 < P ? za[u + 1] : 0.0, acc1);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qq = qt * 16 + 4 * i + kr;
      if (qq < k && kc < B) Cm[qq * B + kc] = acc0[i] + acc1[i];
    }
  }
  ky_barrier();
  const int nt = (P + 15) / 16;
  const double *cm = Cm + (kc < B ? kc : 0);
  for (int ct = wave; ct < nt; ct += kKyWaves) {
    const int c = ct * 16 + kc;
    const bool cin = c < P;
    const double *qc = a.Qt + (cin ? c : P - 1);
    kd4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < k; q0 += 32) {
      double ya[8], ca[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int qq = q0 + 4 * u + kr;
        ya[u] = qc[(size_t)(qq < k ? qq : k - 1) * P];
        ca[u] = cm[(qq < k ? qq : k - 1) * B];
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        const int q1 = q0 + 4 * u + kr, q2 = q1 + 4;
        acc0 = ky_mfma(kin && q1 < k ? ca[u] : 0.0, q1 < k && cin ? ya[u] : 0.0, acc0);
        acc1 = ky_mfma(kin && q2 < k ? ca[u + 1] : 0.0, q2 < k && cin ? ya[u + 1] : 0.0, acc1);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * i + kr;
      if (r < B && c < P) Zb[r * kKyMaxP + c] -= acc0[i] + acc1[i];
    }
  }
  ky_barrier();
}

// Squared norms of the rows r < b of Zb into nrm (LDS), one wave per row.
template <int B>
__device__ void ky_row_norms(const KyArgs &a, const double *Zb, double *nrm) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int r = wave; r < a.b; r += kKyWaves) {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double z = Zb[r * kKyMaxP + lane + 64 * u];
      s += z * z;
    }
    s = ky_wave_sum(s);
    if (lane == 0) nrm[r] = s;
  }
  ky_barrier();
}

__device__ inline double ky_readlane(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Z = D^{-1/2} L^{-1} Z with Z Z^T = L D L^T (one Cholesky QR pass on the
// rows r < b).  The Gram on the matrix cores (K = the P columns split over
// the waves, the wave partials added in a fixed order), the factors by wave 0
// with row r in the registers of lane r (Mg keeps L below the diagonal and
// D^{-1/2} on it).  Returns false (in every thread) when a
// pivot falls to 1e-20 of ref[r] (squared row norms before the projection)
// or is not positive.
template <int B>
__device__ bool ky_cholqr(const KyArgs &a, double *Zb, double *Mg, double *part,
                          const double *ref, int *flag) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int kr = lane >> 4, kc = lane & 15, P = a.P, b = a.b;
  {
    kd4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const double *zr = Zb + (kc < B ? kc : 0) * kKyMaxP;
    for (int c0 = 4 * wave; c0 < P; c0 += 8 * kKyWaves) {
      const int c1 = c0 + kr, c2 = c1 + 4 * kKyWaves;
      const double z1 = zr[c1 < kKyMaxP ? c1 : 0], z2 = zr[c2 < kKyMaxP ? c2 : 0];
      const double x1 = kc < B && c1 < P ? z1 : 0.0, x2 = kc < B && c2 < P ? z2 : 0.0;
      acc0 = ky_mfma(x1, x1, acc0);
      acc1 = ky_mfma(x2, x2, acc1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) part[(wave * 16 + 4 * i + kr) * 16 + kc] = acc0[i] + acc1[i];
  }
  ky_barrier();
  if (wave == 0) {
    // lane r: row r of the Gram (sum of the wave partials in wave order);
    // rows b .. B-1 are those of the identity, so that the factorisation
    // runs over all B rows without a branch (their pivots are 1, their
    // factor rows those of the identity)
    const int lr = lane < B ? lane : 0;
    double m[B];
#pragma unroll
    for (int s = 0; s < B; ++s) {
      double x = 0.0;
#pragma unroll
      for (int w = 0; w < kKyWaves; ++w) x += part[(w * 16 + lr) * 16 + s];
      m[s] = lane < b ? x : lane == s ? 1.0 : 0.0;
    }
    const double rf = lane < b ? ref[lane] : 0.0;
    // M = L D L^T with L unit lower triangular (no square root on the
    // dependent chain; the pivots' reciprocals by v_rcp_f64 and two Newton
    // steps, each squaring the relative error)
    double dl = 1.0;  // lane r: D[r]
    bool good = true;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const double d = ky_readlane(m[j], j);
      good = good && (j >= b || d > 1e-20 * ky_readlane(rf, j));
      const double dd = good ? d : 1.0;
      double inv = __builtin_amdgcn_rcp(dd);
      inv = fma(inv, fma(-dd, inv, 1.0), inv);
      inv = fma(inv, fma(-dd, inv, 1.0), inv);
      if (lane == j) dl = dd;
      double u[B];
#pragma unroll
      for (int s = j + 1; s < B; ++s) u[s] = ky_readlane(m[j], s);
      const double lrj = m[j] * inv;
#pragma unroll
      for (int s = j + 1; s < B; ++s)
        if (lane >= s) m[s] -= lrj * u[s];
      if (lane > j) m[j] = lrj;
    }
    const double rs = 1.0 / sqrt(dl);
    const int ok = good ? 1 : 0;
    if (lane < B) {
#pragma unroll
      for (int s = 0; s < B; ++s) Mg[lane * B + s] = lane == s ? rs : m[s];
    }
    if (lane == 0) *flag = ok;
  }
  ky_barrier();
  const bool ok = *flag != 0;
  // forward substitution by columns
  const int c = t;
  if (ok && c < P) {
    double z[B];
#pragma unroll
    for (int r = 0; r < B; ++r) z[r] = Zb[r * kKyMaxP + c];
#pragma unroll
    for (int r = 0; r < B; ++r) {
      double x = z[r];
#pragma unroll
      for (int s = 0; s < r; ++s) x -= Mg[r * B + s] * z[s];
      z[r] = x;
    }
#pragma unroll
    for (int r = 0; r < B; ++r) z[r] *= Mg[r * B + r];
#pragma unroll
    for (int r = 0; r < B; ++r) Zb[r * kKyMaxP + c] = z[r];
  }
  ky_barrier();
  return ok;
}

template <int B>
__global__ __launch_bounds__(kKyThreads) void krylov_basis_kernel(KyArgs a) {
  extern __shared__ double lds[];
  double *Zb = lds;                  // [B][256] the current block, rows
  double *QC = Zb + B * kKyMaxP;     // [256][B] the last orthonormal block transposed
                                     // (product with G), then projection coefficients
  double *part = QC + kKyMaxP * B;   // [waves][16][16] Gram partials
  double *Mg = part + kKyWaves * 256;  // [B][B] Cholesky factor
  double *ref = Mg + B * B;            // [B] squared row norms before projecting
  int *flag = reinterpret_cast<int *>(ref + B);
  const int t = threadIdx.x, P = a.P, b = a.b, m = a.m;
  for (int x = t; x < B * kKyMaxP; x += kKyThreads) {
    const int r = x / kKyMaxP, c = x % kKyMaxP;
    Zb[x] = r < b && c < P ? a.X0t[(size_t)r * P + c] : 0.0;
  }
  ky_barrier();
  ky_row_norms<B>(a, Zb, ref);
  bool ok = ky_cholqr<B>(a, Zb, Mg, part, ref, flag);
  if (ok) {
    for (int x = t; x < b; x += kKyThreads) ref[x] = 1.0;
    ky_barrier();
    ok = ky_cholqr<B>(a, Zb, Mg, part, ref, flag);
  }
  int built = 0;  // blocks with Q_j and (Q G)_j written
  for (int j = 0; ok && j < m; ++j) {
    // Q_j to the output rows and, transposed, for the product with G; the
    // projections read the output rows back (other threads' stores: a
    // barrier with a workgroup fence on global memory)
    for (int x = t; x < B * kKyMaxP; x += kKyThreads) {
      const int r = x / kKyMaxP, c = x % kKyMaxP;
      const double z = Zb[x];
      QC[c * B + r] = r < b ? z : 0.0;
      if (r < b && c < P) a.Qt[(size_t)(j * b + r) * P + c] = z;
    }
    __syncthreads();
    ky_gprod<B>(a, j * b, Zb, QC);
    built = j + 1;
    if (j + 1 == m) break;
    // project, normalise, and again: the second projection comes after the
    // first normalisation, which scales up whatever of the earlier blocks the
    // first left in directions of small norm
    ky_row_norms<B>(a, Zb, ref);
    ky_project<B>(a, (j + 1) * b, Zb, QC);
    ok = ky_cholqr<B>(a, Zb, Mg, part, ref, flag);
    if (ok) {
      ky_project<B>(a, (j + 1) * b, Zb, QC);
      for (int x = t; x < b; x += kKyThreads) ref[x] = 1.0;
      ky_barrier();
      ok = ky_cholqr<B>(a, Zb, Mg, part, ref, flag);
    }
  }
  // rows past the last block built: zero (no other store reaches them)
  for (size_t x = (size_t)built * b * P + t; x < (size_t)m * b * P; x += kKyThreads) {
    a.Qt[x] = 0.0;
    a.GQt[x] = 0.0;
  }
}

template <int B>
size_t ky_lds() {
  return (size_t)(2 * B * kKyMaxP + kKyWaves * 256 + B * B + B + 1) * sizeof(double);
}

}  // namespace

bool krylov_basis(Ctx &c, const double *G, const double *X0t, u64 P, u32 b, u32 m, double *Qt,
                  double *GQt) {
  if (P < 1 || P > (u64)kKyMaxP || b < 1 || b > 16 || m < 1 || (u64)m * b > P) return false;
  KyArgs a{G, X0t, Qt, GQt, (int)P, (int)b, (int)m};
  // block widths 4, 8, 12, 16: the factorisation of each block is quadratic
  // in the width, padded rows included
  auto launch = [&](auto kern, size_t lds) {
    allow_dyn_lds(c.device, (const void *)kern, lds);
    kern<<<1, kKyThreads, lds, c.stream>>>(a);
  };
  if (b <= 4) launch(krylov_basis_kernel<4>, ky_lds<4>());
  else if (b <= 8) launch(krylov_basis_kernel<8>, ky_lds<8>());
  else if (b <= 12) launch(krylov_basis_kernel<12>, ky_lds<12>());
  else launch(krylov_basis_kernel<16>, ky_lds<16>());
  SP_CHECK_LAUNCH();
  return true;
}

}  // namespace spamd
