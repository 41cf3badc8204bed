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
  hipFuncAttributes fa;
  SP_HIP(hipFuncGetAttributes(&fa, kern));
  // reflectors per LDS stage: as many as fit, at most 32 (one tau each)
  const size_t budget = kTrdLds - fa.sharedSizeBytes;
  const int B = (int)std::min<u64>(32, budget / ((P + 1) * sizeof(double)));
  if (B < 1) return false;
  const size_t lds = (size_t)B * (P + 1) * sizeof(double);
  SP_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
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
  const int per = P <= (u64)kTrdThreads       ? 1
                  : P <= 2 * (u64)kTrdThreads ? 2
                  : P <= 4 * (u64)kTrdThreads ? 4
                                              : 8;
  const void *kern = per == 1   ? (const void *)sytrd_lower_kernel<1>
                     : per == 2 ? (const void *)sytrd_lower_kernel<2>
                     : per == 4 ? (const void *)sytrd_lower_kernel<4>
                                : (const void *)sytrd_lower_kernel<8>;
  int ncu = 0;
  SP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
  // the LDS a workgroup may take: 160 KiB less the kernel's static part
  hipFuncAttributes fa;
  SP_HIP(hipFuncGetAttributes(&fa, kern));
  const size_t budget = kTrdLds - fa.sharedSizeBytes;
  // about SP_SYTRD_ROWS rows per workgroup, more workgroups when the rows do
  // not fit the LDS; one workgroup per CU at most (co-residency) and at most
  // one share of p.v per thread of the final sum
  const int wmax = std::min(ncu, kTrdThreads);
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
  SP_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
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
