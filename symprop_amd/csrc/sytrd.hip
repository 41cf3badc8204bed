// Reduction of the symmetric P x P Gram to tridiagonal form, G = Q T Q^T,
// in one launch of co-resident workgroups (the eigen-step's first stage; the
// step follows matrix::syevd, src/dense/matrixops.cpp:77-100).
//
// Arithmetic of LAPACK dsytd2 with uplo = 'L': for column k, the reflector
// H_k = I - tau_k v v^T (v_{k+1} = 1) of dlarfg annihilates A[k+2:, k]; then
// p = tau_k A22 v, w = p - (tau_k/2)(p.v) v and A22 -= v w^T + w v^T.  The
// output layout is LAPACK's (and rocSOLVER's) for uplo = lower, so that
// ormtr applies Q: d (P), e (P-1), tau (P-1), v_k's tail in column k of a
// P x P array below the subdiagonal (column-major).  G itself is only read.
//
// Layout: row i of G lives in the LDS of workgroup i mod W for the whole
// reduction (full rows, so a row's symv and rank-2 update read only local
// memory).  Per column there is one exchange through global memory: every
// workgroup publishes its rows of p, its share of p.v, and (the owner of row
// k+1) row k+1 as it stands, then counts itself in and waits for the rest.
// Each workgroup then applies the update to row k+1 itself and derives the next
// reflector from it, with the same order of operations everywhere, so no
// second exchange is needed to share v.  All reductions have a fixed order,
// so the result is deterministic.  Each wait is bounded: on a time-out the
// kernel raises a flag, every workgroup leaves at its next wait, and the host
// reports that the reduction was not done (G is never written).
#include "sp_internal.hpp"
#include "wave.hpp"

#include <algorithm>

// build-time tuning: rows of G per workgroup (target) and threads per
// workgroup of the reduction
#ifndef SP_SYTRD_ROWS
#define SP_SYTRD_ROWS 16
#endif
#ifndef SP_TRD_THREADS
#define SP_TRD_THREADS 512
#endif

namespace spamd {

namespace {

constexpr int kTrdThreads = SP_TRD_THREADS;
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kTrdMaxPer = 8;  // row elements per thread: P <= 2048
constexpr int kTrdGroup = 4;   // rows per pass of the update loop
constexpr int kSymvGroup = kTrdWaves >= 8 ? 2 : 4;  // rows per wave per pass of the symv
constexpr size_t kTrdLds = 160 * 1024;  // LDS of a CU

struct TrdArgs {
  const double *G;  // the P x P matrix, read only
  double *Vr;       // reflector tails: v_k's rows k+2.. in column k (column-major)
  int P, W, nloc;
  double *d, *e, *tau;
  double *pbuf;     // [2][P]: p of the column (by column parity)
  double *xbuf;     // [2][P]: row k+1 before column k's update (by parity)
  double *part;     // [2][W]: each workgroup's share of p.v (by parity)
  unsigned *ctr;    // arrivals at the exchanges (zeroed before the launch)
  int *abort;       // raised by a wait that timed out
  unsigned spin_limit;
};

// Sum over the workgroup in a fixed order; every thread gets the result.
__device__ inline double block_sum(double x, double *red) {
  x = wave_sum(x);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < kTrdWaves; ++i) s += red[i];
  __syncthreads();
  return s;
}

// dlarfg on x[k+1:] (x = row k of the partly reduced matrix): v (v[k+1] = 1)
// into LDS, tau and beta returned.  Every workgroup runs this on the same
// row with the same order of operations, so all hold the same reflector.
__device__ inline void reflector(const double *x, int k, int P, double *v, double *red,
                                 double &tk, double &beta) {
  double ss = 0.0;
  for (int j = k + 2 + (int)threadIdx.x; j < P; j += kTrdThreads) ss += x[j] * x[j];
  ss = block_sum(ss, red);
  const double alpha = x[k + 1];
  double scal = 0.0;
  beta = alpha;
  tk = 0.0;
  if (ss > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    tk = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  for (int j = k + 1 + (int)threadIdx.x; j < P; j += kTrdThreads)
    v[j] = j == k + 1 ? 1.0 : x[j] * scal;
  __syncthreads();
}

// Waits until this wave's stores have been acknowledged (s_waitcnt
// vmcnt(0); expcnt and lgkmcnt left at their maxima), i.e. have reached the
// L2, the level the workgroup's single release below writes back.
__device__ inline void stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Publishes this workgroup's data of column `epoch - 1` and waits until every
// workgroup has done so.  One agent-scope release and one acquire per
// workgroup (not per wave): each of these is an L2-wide writeback or
// invalidate, and per wave they cost more than the exchange itself.
// Returns false when the wait timed out (here or in another workgroup).
__device__ bool exchange(const TrdArgs &a, unsigned epoch, int *fail) {
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicAdd(a.ctr, 1u);
    const unsigned target = epoch * (unsigned)a.W;
    int f = 0;
    for (unsigned n = 1;
         __hip_atomic_load(a.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++n) {
      if (n >= a.spin_limit) {
        atomicExch(a.abort, 1);
        f = 1;
        break;
      }
      if ((n & 63) == 0 &&
          __hip_atomic_load(a.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        f = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *fail = f;
  }
  __syncthreads();
  return *fail == 0;
}

// p_i = tau (A v)_i for the G local rows s, s+waves, ... (one wave; lanes over
// the columns j >= k+1, v read once per column for all G rows); stores p_i
// and returns sum p_i v_i in row order.
template <int G>
__device__ inline double symv_rows(const double *A, int P, int W, int wg, int s, int k,
                                   const double *v, double tk, int lane, double *pb) {
  const double *row[G];
  double acc[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    row[g] = A + (size_t)(s + g * kTrdWaves) * P;
    acc[g] = 0.0;
  }
  for (int j = k + 1 + lane; j < P; j += 64) {
    const double vj = v[j];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] += row[g][j] * vj;
  }
  double pv = 0.0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int i = wg + (s + g * kTrdWaves) * W;
    const double p = tk * wave_sum(acc[g]);
    pv += p * v[i];
    if (lane == 0) pb[i] = p;
  }
  return pv;
}

// A[i][j] -= v_i w_j + w_i v_j for the G local rows s..s+G-1 and this
// thread's columns (v_j, w_j in rv, rw); all loads of the group are issued
// before its first store.
template <int G, int PER>
__device__ inline void update_rows(double *A, int P, int W, int wg, int s, int k, int t,
                                   const double *v, const double *w, const double *rv,
                                   const double *rw) {
  double r[G][PER], vi[G], wi[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const double *row = A + (size_t)(s + g) * P;
    vi[g] = v[wg + (s + g) * W];
    wi[g] = w[wg + (s + g) * W];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = k + 1 + t + u * kTrdThreads;
      r[g][u] = j < P ? row[j] : 0.0;
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    double *row = A + (size_t)(s + g) * P;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = k + 1 + t + u * kTrdThreads;
      if (j < P) row[j] = r[g][u] - (vi[g] * rw[u] + wi[g] * rv[u]);
    }
  }
}

// One exchange per column: after column k's exchange every workgroup holds
// all of p, and row k+1 as it was before column k's update; each applies the
// update to that row and to its own rows, and derives the next reflector
// from the updated row itself.
//
// PER: row elements per thread (P <= kTrdThreads*PER).  Each thread owns
// columns j = k+1+t+kTrdThreads*u, u < PER, of every row it touches outside
// the symv; the row loops take several rows at a time while that many
// remain, so the LDS round trips of a group overlap.
template <int PER>
__global__ __launch_bounds__(kTrdThreads) void sytrd_lower_kernel(TrdArgs a) {
  extern __shared__ double lds[];
  __shared__ double red[kTrdWaves];
  __shared__ int fail;
  const int P = a.P, W = a.W, wg = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  double *A = lds;                      // local row s is global row wg + s*W
  double *v = A + (size_t)a.nloc * P;  // [P] reflector of the column
  double *w = v + P;                    // [P]
  double *x = w + P;                    // [P] pivot row
  double *rc = x + P;                   // [waves] p_i v_i of each wave's rows
  const int nrows = wg < P ? (P - wg + W - 1) / W : 0;
  for (int s = 0; s < nrows; ++s) {
    const double *src = a.G + (size_t)(wg + s * W) * P;
    for (int j = t; j < P; j += kTrdThreads) A[(size_t)s * P + j] = src[j];
  }
  for (int j = t; j < P; j += kTrdThreads) x[j] = a.G[j];
  __syncthreads();
  for (int k = 0; k + 1 < P; ++k) {
    const int par = k & 1;
    double tk, beta;
    reflector(x, k, P, v, red, tk, beta);
    // p = tau A v over the local rows i >= k+1: wave q takes rows s0+q,
    // s0+q+waves, ..., kSymvGroup of them per pass over v
    const int s0 = k + 1 <= wg ? 0 : (k + 1 - wg + W - 1) / W;
    double *pb = a.pbuf + (size_t)par * P;
    double pv = 0.0;  // this wave's rows' p_i v_i, in row order
    int sq = s0 + wave;
    for (; sq + (kSymvGroup - 1) * kTrdWaves < nrows; sq += kSymvGroup * kTrdWaves)
      pv += symv_rows<kSymvGroup>(A, P, W, wg, sq, k, v, tk, lane, pb);
    for (; sq < nrows; sq += kTrdWaves) pv += symv_rows<1>(A, P, W, wg, sq, k, v, tk, lane, pb);
    if (lane == 0) rc[wave] = pv;
    double *xb = a.xbuf + (size_t)par * P;
    if ((k + 1) % W == wg) {
      const double *row = A + (size_t)((k + 1) / W) * P;
      for (int j = k + 1 + t; j < P; j += kTrdThreads) xb[j] = row[j];
    }
    __syncthreads();
    if (t == 0) {
      double c = 0.0;
      for (int q = 0; q < kTrdWaves; ++q) c += rc[q];
      a.part[(size_t)par * W + wg] = c;
    }
    if (!exchange(a, (unsigned)k + 1, &fail)) return;
    // gather p, the pivot row and the shares of p.v: every load of this
    // thread is issued before the first is used (one memory round trip)
    double rp[PER], rx[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = k + 1 + t + u * kTrdThreads;
      rp[u] = j < P ? pb[j] : 0.0;
      rx[u] = j < P ? xb[j] : 0.0;
    }
    const double share = t < W ? a.part[(size_t)par * W + t] : 0.0;
    if (k % W == wg) {  // column k of the output
      for (int j = k + 2 + t; j < P; j += kTrdThreads) a.Vr[(size_t)k * P + j] = v[j];
      if (t == 0) {
        a.d[k] = x[k];
        a.e[k] = beta;
        a.tau[k] = tk;
      }
    }
    const double dot = block_sum(share, red);
    const double al = -0.5 * tk * dot;
    double rv[PER], rw[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = k + 1 + t + u * kTrdThreads;
      rv[u] = j < P ? v[j] : 0.0;
      rw[u] = rp[u] + al * rv[u];
      if (j < P) w[j] = rw[u];
    }
    __syncthreads();
    {
      const double vk = v[k + 1], wk = w[k + 1];
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int j = k + 1 + t + u * kTrdThreads;
        if (j < P) x[j] = rx[u] - (vk * rw[u] + wk * rv[u]);
      }
    }
    int su = s0;
    for (; su + kTrdGroup - 1 < nrows; su += kTrdGroup)
      update_rows<kTrdGroup, PER>(A, P, W, wg, su, k, t, v, w, rv, rw);
    for (; su < nrows; ++su) update_rows<1, PER>(A, P, W, wg, su, k, t, v, w, rv, rw);
    __syncthreads();
  }
  if ((P - 1) % W == wg && t == 0) a.d[P - 1] = x[P - 1];
}

// ---------------------------------------------------------------------------
// One-workgroup form for P <= 256: the same arithmetic (dsytd2, uplo = 'L')
// and output layout, with no exchange through global memory.
//
// The lower triangle of G, cut into 16 x 16 blocks (diagonal blocks stored
// whole), lives in the registers of 512 threads: wave q holds list entries
// q, q+8, q+16, ... of the blocks ordered by block column, last first, so the
// blocks still active at column k (block column >= (k+1)/16) are a prefix of
// every wave's list and the work stays balanced as the trailing block
// shrinks.  In a block, lane l holds row l%16 and columns 4(l/16)..+3.
//
// Per column k, three phases and three workgroup barriers:
//   A  every wave applies column k-1's rank-2 update to its blocks and forms
//      the partial products of p = tau_k A v_k: block (I,J) gives its row
//      sums to S[I][J] and, off the diagonal, its column sums to S[J][I];
//      the holders of column k+1 copy it, as updated so far, to the LDS;
//   B  wave q sums S[I][J] over J (fixed order) into p for the block rows
//      I = Jk+q, Jk+q+8;
//   C  wave 0 forms w = p - (tau/2)(p.v) v, applies the update to column
//      k+1, derives the next reflector from it (dlarfg) and writes w, the new
//      v and column k+1's d, e, tau to the LDS; four waves copy them (and
//      v's tail) to the outputs during the next phase A.
// The other waves wait at the third barrier: phase C is one dependent chain
// (two wave sums, a square root, a division), and run by all waves at once it
// would only take issue slots from each other.  Deterministic: every sum has
// a fixed order, and the wave sums give every lane the same bits.
//
// Measured at P = 220 (cycles per column): phase A ~4,000, phase B ~770,
// phase C ~2,000, barrier waits ~1,000; 0.69 ms against 0.83 ms for the
// multi-workgroup form, 0.29 against 0.46 ms at P = 130.
constexpr int kOwThreads = 512;
constexpr int kOwWaves = kOwThreads / 64;
constexpr int kOwNB = 16;  // block rows: P <= 256
constexpr int kOwMaxP = 16 * kOwNB;
constexpr int kOwBlocks = kOwNB * (kOwNB + 1) / 2;
constexpr int kOwSlots = (kOwBlocks + kOwWaves - 1) / kOwWaves;  // blocks per wave
constexpr int kOwSpare = kOwNB * kOwNB * 16;                    // spare row of S
// LDS (doubles): S [NB][NB][16] + spare row, column k+1 [256] + spare row,
// v [2][256], w [256], p [256], d/e/tau [2][3]
constexpr size_t kOwLds = (size_t)(kOwSpare + 16 + 5 * kOwMaxP + 64 + 8) * sizeof(double);

struct OwArgs {
  const double *G;
  double *Vr;
  int P;
  double *d, *e, *tau;
};

// Workgroup barrier ordering the LDS only (the global stores of the outputs
// need no ordering inside the kernel).
__device__ inline void ow_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// x summed with the same lane of the other half-row pair (lane ^ 16), in the
// same order on both sides: v_permlane16_swap gives each lane of rows 0/1
// (2/3) row 0's and row 1's (row 2's and row 3's) value.
__device__ inline double ow_sum16(double x) {
  const long long b = __double_as_longlong(x);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi =
      __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double a0 = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
  const double a1 = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
  return a0 + a1;
}

// x summed with lane ^ 32 (v_permlane32_swap: lower and upper half-wave).
__device__ inline double ow_sum32(double x) {
  const long long b = __double_as_longlong(x);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double a0 = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
  const double a1 = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
  return a0 + a1;
}

// One DPP move of a double within rows of 16 lanes (every source lane valid).
template <int C>
__device__ inline double ow_dpp(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, C, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), C, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the wave, the same bits in every lane: pairs that swap values
// (lane ^ 1, ^ 2 in the quad, then the half-row and row mirrors, which pair
// lanes whose partial sums are already equal within their quad / half-row,
// then the two half-row and the two half-wave swaps) each add the same two
// operands.
__device__ inline double ow_wave_sum(double x) {
  x += ow_dpp<0xb1>(x);
  x += ow_dpp<0x4e>(x);
  x += ow_dpp<0x141>(x);
  x += ow_dpp<0x140>(x);
  return ow_sum32(ow_sum16(x));
}

// Entry idx (uniform) of a vector held as x[u] at index lane + 64u.
__device__ inline double ow_at(const double x[4], int idx) {
  const int u = idx >> 6;
  const double s = u == 0 ? x[0] : u == 1 ? x[1] : u == 2 ? x[2] : x[3];
  return read_lane(s, idx & 63);
}

// dlarfg on x[kp:] (x at index lane + 64u, entries above kp ignored): v with
// v[kp] = 1 and zeros elsewhere above, tau and beta.
__device__ inline void ow_reflector(const double x[4], int kp, int P, int lane, double v[4],
                                    double &tk, double &beta) {
  double ss = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int j = lane + 64 * u;
    if (j > kp && j < P) ss += x[u] * x[u];
  }
  ss = ow_wave_sum(ss);
  const double alpha = ow_at(x, kp);
  double scal = 0.0;
  beta = alpha;
  tk = 0.0;
  if (ss > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    tk = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int j = lane + 64 * u;
    v[u] = j == kp ? 1.0 : j > kp && j < P ? x[u] * scal : 0.0;
  }
}

// Phase A on one block (I, J) held in a[4]: column k-1's update with vo, wo
// (zero for k = 0), the row and column sums of a against vn into S, and
// column k+1's entries into cbk.  Stores with nothing to store go to spare
// LDS rows instead of branching (`on` false sends them all there).
__device__ inline void ow_block(double a[4], int I, int J, bool on, int Jk, int kc, int lane,
                                const double *vo, const double *wo, const double *vn,
                                double *S, double *cbk) {
  const int r = lane & 15, cg = lane >> 4;
  const int i = I * 16 + r, j0 = J * 16 + 4 * cg;
  const double voi = vo[i], woi = wo[i], vni = vn[i];
  double voj[4], woj[4], vnj[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    voj[q] = vo[j0 + q];
    woj[q] = wo[j0 + q];
    vnj[q] = vn[j0 + q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) a[q] = fma(-woi, voj[q], fma(-voi, woj[q], a[q]));
  // column k+1 (block column Jk, column kc of the block), as updated so far
  const bool mine = on && J == Jk && cg == (kc >> 2);
  const int qc = kc & 3;
  cbk[mine ? i : kOwMaxP + lane] =
      qc == 0 ? a[0] : qc == 1 ? a[1] : qc == 2 ? a[2] : a[3];
  double rs = a[0] * vnj[0];
#pragma unroll
  for (int q = 1; q < 4; ++q) rs += a[q] * vnj[q];
  rs = ow_sum32(ow_sum16(rs));  // the four column groups, same value in each
  S[on ? (I * kOwNB + J) * 16 + r : kOwSpare + r] = rs;
  // column sums over the block's 16 rows (the lanes of a DPP row): halves
  // swapped with the mirrored lane (bit 3 differs), then with the half-row
  // mirror (bit 2 differs), then summed over the quad; the quad then holds
  // column 2*bit3 + bit2 of its group.  The diagonal block's rows already
  // cover both halves.
  const double c0 = a[0] * vni, c1 = a[1] * vni, c2 = a[2] * vni, c3 = a[3] * vni;
  const bool h8 = lane & 8, h4 = lane & 4;
  double k0 = h8 ? c2 : c0, k1 = h8 ? c3 : c1;
  k0 += ow_dpp<0x140>(h8 ? c0 : c2);
  k1 += ow_dpp<0x140>(h8 ? c1 : c3);
  double y = h4 ? k1 : k0;
  y += ow_dpp<0x141>(h4 ? k0 : k1);
  y += ow_dpp<0xb1>(y);
  y += ow_dpp<0x4e>(y);
  S[on && I != J ? (J * kOwNB + I) * 16 + 4 * cg + (h8 ? 2 : 0) + (h4 ? 1 : 0) : kOwSpare + r] = y;
}

__global__ __launch_bounds__(kOwThreads) void sytrd_ow_kernel(OwArgs a) {
  extern __shared__ double lds[];
  double *S = lds;                   // [I][J][16] partial products of p, then a spare row
  double *cb = S + kOwSpare + 16;    // column k+1 [256], then a spare row of 64
  double *vb = cb + kOwMaxP + 64;    // [2][256] v, by parity of its column
  double *wb = vb + 2 * kOwMaxP;     // [256] w
  double *pb = wb + kOwMaxP;         // [256] p
  double *sc = pb + kOwMaxP;         // [2][3] d, e, tau of a column, by parity
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int P = a.P, NB = (P + 15) >> 4, nblk = NB * (NB + 1) / 2;
  const int r = lane & 15, cg = lane >> 4;

  // this wave's blocks: list entry m = s*waves + wave lies on diagonal d of
  // the list (block column NB-1-d) at offset o (block row NB-1-o); the
  // coordinates (I << 4 | J) of four slots are packed per word
  unsigned pk[(kOwSlots + 3) / 4] = {};
  double A[kOwSlots][4];
#pragma unroll
  for (int s = 0; s < kOwSlots; ++s) {
    const int m = s * kOwWaves + wave;
    const bool on = m < nblk;
    int d = 0;
    while ((d + 1) * (d + 2) / 2 <= m) ++d;
    const int I = on ? NB - 1 - (m - d * (d + 1) / 2) : 0, J = on ? NB - 1 - d : 0;
    pk[s / 4] |= (unsigned)(I << 4 | J) << (8 * (s % 4));
    const int i = I * 16 + r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = J * 16 + 4 * cg + q;
      // lower triangle, column-major (its transpose above the diagonal)
      A[s][q] = on && i < P && j < P
                    ? (i >= j ? a.G[(size_t)j * P + i] : a.G[(size_t)i * P + j])
                    : 0.0;
    }
  }

  // wave 0's state: v of the current column and its tau
  double vk[4] = {0.0, 0.0, 0.0, 0.0}, tk = 0.0;
  if (wave == 0) {
    double x[4], beta;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = lane + 64 * u;
      x[u] = j < P ? a.G[j] : 0.0;
    }
    ow_reflector(x, 1, P, lane, vk, tk, beta);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = lane + 64 * u;
      vb[j] = vk[u];
      vb[kOwMaxP + j] = 0.0;
      wb[j] = 0.0;
      if (j >= 2 && j < P) a.Vr[j] = vk[u];
    }
    if (lane == 0) {
      a.d[0] = x[0];
      a.e[0] = beta;
      a.tau[0] = tk;
    }
  }
  ow_barrier();

  for (int k = 0; k + 1 < P; ++k) {
    const int par = k & 1;
    const int Jk = (k + 1) >> 4, kc = (k + 1) & 15;
    const int na = (NB - Jk) * (NB - Jk + 1) / 2;  // active blocks
    const double *vo = vb + (par ^ 1) * kOwMaxP;   // v of column k-1 (zero for k = 0)
    const double *vn = vb + par * kOwMaxP;         // v of column k
    if (k > 0 && wave < 4) {
      // column k's outputs (phase C of column k-1 left them in the LDS), a
      // quarter of v's tail per wave
      const int j = lane + 64 * wave;
      if (j > k + 1 && j < P) a.Vr[(size_t)k * P + j] = vn[j];
      if (wave == 0 && lane == 0) {
        const double *o = sc + 3 * par;
        a.d[k] = o[0];
        a.e[k] = o[1];
        a.tau[k] = o[2];
      }
    }
    // phase A
#pragma unroll
    for (int s = 0; s < kOwSlots; ++s) {
      if (s * kOwWaves + wave >= na) break;
      const unsigned c = (pk[s / 4] >> (8 * (s % 4))) & 0xff;
      ow_block(A[s], (int)(c >> 4), (int)(c & 15), true, Jk, kc, lane, vo, wb, vn, S, cb);
    }
    ow_barrier();
    // phase B: the partial products of block rows Jk+wave, Jk+wave+8, four
    // lanes per row (J = Jk+cg, Jk+cg+4, ...), then summed over the four
    {
      double acc[2] = {0.0, 0.0};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int I = Jk + wave + kOwWaves * h;
        if (I < NB)
          for (int J = Jk + cg; J < NB; J += 4) acc[h] += S[(I * kOwNB + J) * 16 + r];
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int I = Jk + wave + kOwWaves * h;
        const double x = ow_sum32(ow_sum16(acc[h]));
        if (I < NB && cg == 0) pb[I * 16 + r] = x;
      }
    }
    ow_barrier();
    // phase C (wave 0); v of column k is 1 at k+1, so w and the updated
    // column there and at k+2 follow from entries every lane reads
    if (wave == 0) {
      const bool more = k + 2 < P;  // column k+1 has a reflector
      const double pk1 = tk * pb[k + 1], ck1 = cb[k + 1];
      const double pk2 = more ? tk * pb[k + 2] : 0.0, ck2 = more ? cb[k + 2] : 0.0;
      const double vk2 = more ? vn[k + 2] : 0.0;
      double p[4], xc[4], w[4], pv = 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = lane + 64 * u;
        const bool live = j > k && j < P;
        p[u] = live ? tk * pb[j] : 0.0;
        xc[u] = live ? cb[j] : 0.0;
        pv += p[u] * vk[u];
      }
      const double al = -0.5 * tk * ow_wave_sum(pv);
      const double wk1 = pk1 + al;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = lane + 64 * u;
        w[u] = p[u] + al * vk[u];
        xc[u] = j > k && j < P ? xc[u] - (vk[u] * wk1 + w[u]) : 0.0;
        wb[j] = w[u];
      }
      double vx[4] = {0.0, 0.0, 0.0, 0.0}, tn = 0.0, bn = 0.0;
      if (more) {
        // dlarfg on the updated column k+1 below k+1
        const double alpha = ck2 - (vk2 * wk1 + (pk2 + al * vk2));
        double ss = 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = lane + 64 * u;
          if (j > k + 2) ss += xc[u] * xc[u];
        }
        ss = ow_wave_sum(ss);
        double scal = 0.0;
        bn = alpha;
        if (ss > 0.0) {
          bn = -copysign(sqrt(alpha * alpha + ss), alpha);
          tn = (bn - alpha) / bn;
          scal = 1.0 / (alpha - bn);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = lane + 64 * u;
          vx[u] = j == k + 2 ? 1.0 : j > k + 2 ? xc[u] * scal : 0.0;
          vb[(par ^ 1) * kOwMaxP + j] = vx[u];
        }
      }
      // column k+1's d, e, tau for the writer of the outputs (d alone for the last)
      if (lane == 0) {
        double *o = sc + 3 * (par ^ 1);
        o[0] = ck1 - (wk1 + wk1);
        o[1] = bn;
        o[2] = tn;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) vk[u] = vx[u];
      tk = tn;
    }
    ow_barrier();
  }
  if (wave == 0 && lane == 0) a.d[P - 1] = sc[3 * ((P - 1) & 1)];
}

// ---------------------------------------------------------------------------
// Multi-workgroup form with the rows in registers (P > 1,024).  The same
// exchange as sytrd_lower_kernel, but each workgroup's rows (at most 16:
// rows wg, wg+W, ...) live in the registers of its 512 threads, thread t
// holding columns t, t+512, ... of every row, so the symv and the rank-2
// update read no memory but the vectors.  A row's sum over the workgroup:
// the 16 rows' partial sums reduced over each wave at once (half-wave and
// half-row swaps, DPP mirrors, quad swaps: 16 values in about 60 VALU
// instructions, lane l ending with row (l/4) mod 16), then over the waves
// through the LDS in a fixed order.
#ifndef SP_SYTRD_REG_ROWS
#define SP_SYTRD_REG_ROWS 16
#endif
constexpr int kRgRows = SP_SYTRD_REG_ROWS;  // rows per workgroup at most (a multiple of 16)

// (a' + b') of v_permlane32_swap(a, b) / v_permlane16_swap(a, b): the lanes
// of the lower half-wave (even half-rows) get their a plus the upper (odd)
// partner's a, the others their b plus the partner's b.
__device__ inline double rg_pair32(double a, double b) {
  const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)y, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((unsigned)(x >> 32), (unsigned)(y >> 32), false, false);
  return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) +
         __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
__device__ inline double rg_pair16(double a, double b) {
  const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)y, false, false);
  const auto hi =
      __builtin_amdgcn_permlane16_swap((unsigned)(x >> 32), (unsigned)(y >> 32), false, false);
  return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) +
         __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}

// x[q] summed over the wave; lane l returns the sum of x[(l >> 2) & 15] (the
// four lanes of a quad the same bits).
__device__ inline double rg_reduce16(const double x[16], int lane) {
  double y[8], z[4], w[2];
#pragma unroll
  for (int q = 0; q < 8; ++q) y[q] = rg_pair32(x[q], x[q + 8]);  // value q + 8*bit5
#pragma unroll
  for (int q = 0; q < 4; ++q) z[q] = rg_pair16(y[q], y[q + 4]);  // value q + 4*bit4 + 8*bit5
  const bool h8 = lane & 8, h4 = lane & 4;
#pragma unroll
  for (int q = 0; q < 2; ++q)  // bit 3: row mirror (lane ^ 15)
    w[q] = (h8 ? z[q + 2] : z[q]) + ow_dpp<0x140>(h8 ? z[q] : z[q + 2]);
  double u = (h4 ? w[1] : w[0]) + ow_dpp<0x141>(h4 ? w[0] : w[1]);  // bit 2: half-row mirror
  u += ow_dpp<0x4e>(u);
  u += ow_dpp<0xb1>(u);
  return u;
}

template <int PER, int ROWS>
__global__ __launch_bounds__(kTrdThreads) void sytrd_reg_kernel(TrdArgs a) {
  extern __shared__ double lds[];
  __shared__ double red[kTrdWaves];
  __shared__ int fail;
  const int P = a.P, W = a.W, wg = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  double *v = lds;                    // [P] reflector of the column
  double *w = v + P;                  // [P]
  double *x = w + P;                  // [P] pivot row
  double *rs = x + P;                 // [ROWS][waves] each wave's sums of the local rows
  double *pvs = rs + ROWS * kTrdWaves;  // [ROWS] p_i v_i of the local rows
  const int nrows = wg < P ? (P - wg + W - 1) / W : 0;
  double A[ROWS][PER];
#pragma unroll
  for (int s = 0; s < ROWS; ++s)
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = t + kTrdThreads * u;
      A[s][u] = s < nrows && j < P ? a.G[(size_t)(wg + s * W) * P + j] : 0.0;
    }
  for (int j = t; j < P; j += kTrdThreads) x[j] = a.G[j];
  __syncthreads();
  for (int k = 0; k + 1 < P; ++k) {
    const int par = k & 1;
    double tk, beta;
    reflector(x, k, P, v, red, tk, beta);
    // this thread's columns of v (v is stale at and before k)
    double rv[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = t + kTrdThreads * u;
      rv[u] = j > k && j < P ? v[j] : 0.0;
    }
#pragma unroll
    for (int g = 0; g < ROWS / 16; ++g) {
      double ps[16];
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        double acc = A[16 * g + s][0] * rv[0];
#pragma unroll
        for (int u = 1; u < PER; ++u) acc += A[16 * g + s][u] * rv[u];
        ps[s] = acc;
      }
      const double r = rg_reduce16(ps, lane);
      if ((lane & 3) == 0) rs[(16 * g + ((lane >> 2) & 15)) * kTrdWaves + wave] = r;
    }
    double *xb = a.xbuf + (size_t)par * P;
    if ((k + 1) % W == wg) {  // row k+1 as it stands, for every workgroup
      const int s1 = (k + 1) / W;
#pragma unroll
      for (int s = 0; s < ROWS; ++s)
        if (s == s1) {
#pragma unroll
          for (int u = 0; u < PER; ++u) {
            const int j = t + kTrdThreads * u;
            if (j > k && j < P) xb[j] = A[s][u];
          }
        }
    }
    __syncthreads();
    double *pb = a.pbuf + (size_t)par * P;
    if (t < ROWS) {
      const int i = wg + t * W;
      double pv = 0.0;
      if (t < nrows && i > k) {
        double sum = 0.0;
        for (int q = 0; q < kTrdWaves; ++q) sum += rs[t * kTrdWaves + q];
        const double p = tk * sum;
        pb[i] = p;
        pv = p * v[i];
      }
      pvs[t] = pv;
    }
    __syncthreads();
    if (t == 0) {
      double c = 0.0;
      for (int s = 0; s < ROWS; ++s) c += pvs[s];
      a.part[(size_t)par * W + wg] = c;
    }
    if (!exchange(a, (unsigned)k + 1, &fail)) return;
    // gather p, the pivot row and the shares of p.v (one memory round trip)
    double rp[PER], rx[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = t + kTrdThreads * u;
      const bool live = j > k && j < P;
      rp[u] = live ? pb[j] : 0.0;
      rx[u] = live ? xb[j] : 0.0;
    }
    const double share = t < W ? a.part[(size_t)par * W + t] : 0.0;
    if (k % W == wg) {  // column k of the output
      for (int j = k + 2 + t; j < P; j += kTrdThreads) a.Vr[(size_t)k * P + j] = v[j];
      if (t == 0) {
        a.d[k] = x[k];
        a.e[k] = beta;
        a.tau[k] = tk;
      }
    }
    const double dot = block_sum(share, red);
    const double al = -0.5 * tk * dot;
    double rw[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = t + kTrdThreads * u;
      rw[u] = rp[u] + al * rv[u];
      if (j < P) w[j] = rw[u];
    }
    __syncthreads();
    {
      const double vk = v[k + 1], wk = w[k + 1];
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int j = t + kTrdThreads * u;
        if (j > k && j < P) x[j] = rx[u] - (vk * rw[u] + wk * rv[u]);
      }
    }
#pragma unroll
    for (int s = 0; s < ROWS; ++s) {
      const int i = wg + s * W;
      if (s < nrows && i > k) {
        const double vi = v[i], wi = w[i];
#pragma unroll
        for (int u = 0; u < PER; ++u) A[s][u] -= vi * rw[u] + wi * rv[u];
      }
    }
    __syncthreads();
  }
  if ((P - 1) % W == wg && t == 0) a.d[P - 1] = x[P - 1];
}

// C = Q Z for Q = H_0 H_1 ... H_{P-2} of sytrd_lower (LAPACK dormtr, side
// left, uplo lower, no transpose).  Each wave owns one column of Z, held in
// registers (lane l has rows l + 64m, m < M); the reflectors are applied from
// the last to the first, staged B at a time through LDS by the whole
// workgroup (each expanded to a full column: zeros above row k+1, the
// implicit 1 at k+1, the stored tail below), with 32 loads in flight per
// thread.  Z is column-major P x R; the result goes to V, row-major P x R.
constexpr int kAqThreads = 256;  // back-transform: 4 waves, one column of Z each
constexpr int kAqWaves = kAqThreads / 64;

template <int M>
__global__ __launch_bounds__(kAqThreads) void apply_q_lower_kernel(const double *__restrict__ G,
                                                            const double *__restrict__ tau,
                                                            int P, int B,
                                                            const double *__restrict__ Z,
                                                            int R, double *__restrict__ V) {
  constexpr int PER = M / 4;        // row elements per thread when staging
  constexpr int QU = 32 / PER;      // reflectors per staging round
  extern __shared__ double vl[];    // [B][P] reflectors, then [B] their tau
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = blockIdx.x * kAqWaves + wave;
  const bool active = r < R;
  double z[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = lane + 64 * m;
    z[m] = active && i < P ? Z[(size_t)r * P + i] : 0.0;
  }
  for (int kb = P - 2; kb >= 0; kb -= B) {
    const int nb = kb + 1 < B ? kb + 1 : B;  // reflectors kb, kb-1, ..., kb-nb+1
    double *tl = vl + (size_t)B * P;
    if (t < nb) tl[t] = tau[kb - t];
    for (int q0 = 0; q0 < nb; q0 += QU) {
      double tmp[QU][PER];
#pragma unroll
      for (int qq = 0; qq < QU; ++qq) {
        const int q = q0 + qq, k = kb - q;
#pragma unroll
        for (int c = 0; c < PER; ++c) {
          const int i = t + kAqThreads * c;
          tmp[qq][c] = q < nb && i < P ? (i > k + 1 ? G[(size_t)k * P + i] : i == k + 1 ? 1.0 : 0.0)
                                       : 0.0;
        }
      }
#pragma unroll
      for (int qq = 0; qq < QU; ++qq) {
        const int q = q0 + qq;
#pragma unroll
        for (int c = 0; c < PER; ++c) {
          const int i = t + kAqThreads * c;
          if (q < nb && i < P) vl[(size_t)q * P + i] = tmp[qq][c];
        }
      }
    }
    __syncthreads();
    if (active) {
      for (int q = 0; q < nb; ++q) {
        const double *vq = vl + (size_t)q * P;
        double vv[M], y = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int i = lane + 64 * m;
          vv[m] = i < P ? vq[i] : 0.0;
          y += vv[m] * z[m];
        }
        const double ty = tl[q] * wave_sum(y);
#pragma unroll
        for (int m = 0; m < M; ++m) z[m] -= ty * vv[m];
      }
    }
    __syncthreads();
  }
  if (active) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = lane + 64 * m;
      if (i < P) V[(size_t)i * R + r] = z[m];
    }
  }
}

}  // namespace

bool apply_q_lower(Ctx &c, const double *G, const double *tau, u64 P, const double *Z, u32 R,
                   double *V) {
  if (P < 2 || P > 2048 || R == 0) return false;
  const int M = P <= 256 ? 4 : P <= 512 ? 8 : P <= 1024 ? 16 : 32;
  const void *kern = M == 4    ? (const void *)apply_q_lower_kernel<4>
                     : M == 8  ? (const void *)apply_q_lower_kernel<8>
                     : M == 16 ? (const void *)apply_q_lower_kernel<16>
                               : (const void *)apply_q_lower_kernel<32>;
  // reflectors per LDS stage: as many as fit, at most 32 (one tau each)
  const size_t budget = kTrdLds - static_lds(c.device, kern);
  const int B = (int)std::min<u64>(32, budget / ((P + 1) * sizeof(double)));
  if (B < 1) return false;
  const size_t lds = (size_t)B * (P + 1) * sizeof(double);
  allow_dyn_lds(c.device, kern, lds);
  int Pi = (int)P, Ri = (int)R;
  void *args[] = {(void *)&G, (void *)&tau, &Pi, (void *)&B, (void *)&Z, &Ri, (void *)&V};
  SP_HIP(hipLaunchKernel(kern, dim3((R + kAqWaves - 1) / kAqWaves), dim3(kAqThreads), args, lds,
                        c.stream));
  SP_CHECK_LAUNCH();
  return true;
}

bool sytrd_lower(Ctx &c, const double *G, u64 P, double *D, double *E, double *tau, double *Vr,
                 const int **abort_flag) {
  if (P < 2 || P > kTrdMaxP || P > (u64)kTrdThreads * kTrdMaxPer) return false;
  if (P <= (u64)kOwMaxP && c.trd_mode == 0) {
    // one workgroup, no waits between workgroups: nothing can time out, so
    // there is no flag to clear or read
    allow_dyn_lds(c.device, (const void *)sytrd_ow_kernel, kOwLds);
    sytrd_ow_kernel<<<1, kOwThreads, kOwLds, c.stream>>>(OwArgs{G, Vr, (int)P, D, E, tau});
    SP_CHECK_LAUNCH();
    *abort_flag = nullptr;
    return true;
  }
  int ncu = 0;
  SP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
  const int wmax = std::min(ncu, kTrdThreads);
  // rows in registers pay off where a workgroup's rows are long (P = 1,716:
  // 15.5 against 20.6 ms); up to P = 1,024 the LDS form is as fast or faster
  if ((c.trd_mode == 2 || (c.trd_mode == 0 && P > 2 * (u64)kTrdThreads)) &&
      P <= 4 * (u64)kTrdThreads && (P + kRgRows - 1) / kRgRows <= (u64)wmax) {
    // rows in registers: 16 per workgroup at most, one workgroup per CU
    const int per = P <= (u64)kTrdThreads ? 1 : P <= 2 * (u64)kTrdThreads ? 2 : 4;
    const void *kern = per == 1   ? (const void *)sytrd_reg_kernel<1, kRgRows>
                       : per == 2 ? (const void *)sytrd_reg_kernel<2, kRgRows>
                                  : (const void *)sytrd_reg_kernel<4, kRgRows>;
    const int W = (int)((P + kRgRows - 1) / kRgRows);
    const size_t lds = (size_t)(3 * P + kRgRows * kTrdWaves + kRgRows) * sizeof(double);
    c.ws_trd.reserve((4 * P + 2 * W) * sizeof(double) + 2 * sizeof(unsigned));
    double *pbuf = c.ws_trd.as<double>();
    double *xbuf = pbuf + 2 * P;
    double *part = xbuf + 2 * P;
    unsigned *ctr = reinterpret_cast<unsigned *>(part + 2 * W);
    int *abort = reinterpret_cast<int *>(ctr + 1);
    SP_HIP(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned), c.stream));
    allow_dyn_lds(c.device, kern, lds);
    TrdArgs a{G, Vr, (int)P, W, kRgRows, D, E, tau, pbuf, xbuf, part, ctr, abort,
              c.trd_spin_limit};
    void *args[] = {&a};
    SP_HIP(hipLaunchKernel(kern, dim3(W), dim3(kTrdThreads), args, lds, c.stream));
    SP_CHECK_LAUNCH();
    *abort_flag = abort;
    return true;
  }
  const int per = P <= (u64)kTrdThreads       ? 1
                  : P <= 2 * (u64)kTrdThreads ? 2
                  : P <= 4 * (u64)kTrdThreads ? 4
                                              : 8;
  const void *kern = per == 1   ? (const void *)sytrd_lower_kernel<1>
                     : per == 2 ? (const void *)sytrd_lower_kernel<2>
                     : per == 4 ? (const void *)sytrd_lower_kernel<4>
                                : (const void *)sytrd_lower_kernel<8>;
  // the LDS a workgroup may take: 160 KiB less the kernel's static part
  const size_t budget = kTrdLds - static_lds(c.device, kern);
  // about SP_SYTRD_ROWS rows per workgroup, more workgroups when the rows do
  // not fit the LDS; one workgroup per CU at most (co-residency) and at most
  // one share of p.v per thread of the final sum
  int W = (int)std::min<u64>((P + SP_SYTRD_ROWS - 1) / SP_SYTRD_ROWS, (u64)wmax);
  W = std::max(W, 1);
  auto lds_of = [P](int w) {
    const u64 nl = (P + w - 1) / w;
    return (size_t)((nl * P + 3 * P + kTrdWaves) * sizeof(double));
  };
  while (W < wmax && lds_of(W) > budget) ++W;
  const size_t lds = lds_of(W);
  if (lds > budget) return false;
  const int nloc = (int)((P + W - 1) / W);
  // scratch: pbuf (2P), xbuf (2P), part (2W), then the counter and abort words
  c.ws_trd.reserve((4 * P + 2 * W) * sizeof(double) + 2 * sizeof(unsigned));
  double *pbuf = c.ws_trd.as<double>();
  double *xbuf = pbuf + 2 * P;
  double *part = xbuf + 2 * P;
  unsigned *ctr = reinterpret_cast<unsigned *>(part + 2 * W);
  int *abort = reinterpret_cast<int *>(ctr + 1);
  SP_HIP(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned), c.stream));
  allow_dyn_lds(c.device, kern, lds);
  TrdArgs a{G, Vr, (int)P, W, nloc, D, E, tau, pbuf, xbuf, part, ctr, abort, c.trd_spin_limit};
  // An ordinary launch: W <= CUs workgroups, each within one CU's LDS, are
  // resident together on the device.  (A cooperative launch would also
  // reserve the hardware wave-sync unit, which one process at a time can
  // hold; the exchange here is in software and does not need it.)
  void *args[] = {&a};
  SP_HIP(hipLaunchKernel(kern, dim3(W), dim3(kTrdThreads), args, lds, c.stream));
  SP_CHECK_LAUNCH();
  // the caller reads the flag with the kernel's output (one wait for both)
  *abort_flag = abort;
  return true;
}

}  // namespace spamd
