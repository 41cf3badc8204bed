// Fast S3TTMc kernels: compile-time (leaf size M = N-1, rank R, head degree H).
//
// With the polynomial form of the leaf tensor (see s3ttmc.hip), the packed Y
// row of index r is
//     Y[r] = fact o fold( sum_{owners o of r} (c_o * head_o) (x) tail_o ),
// where head_o is the product of the linear forms of the first H leaf rows
// (P_H packed coefficients), tail_o that of the last T = M-H rows (P_T), fold
// adds every (h, t) pair into the packed entry of the merged multi-index and
// fact multiplies by prod(f_J!).  The owner sum is a GEMM with K = owners:
//     Ypre (P_H x P_T) = A^T B,   A = [c_o head_o],  B = [tail_o],
// run on FP64 MFMA (v_mfma_f64_16x16x4f64, 16x16 tiles, K = 4 owners).
//
// Work split of a 256-thread workgroup over a batch of 64 (or 48) owners of one row:
//   producer: lane = owner, wave w computes the w-th quarter of the head and
//             tail coefficients in registers (fully unrolled, constant
//             indices) and stores them to LDS;
//   consumer: wave w owns output tiles w, w+4, ...; each MFMA reads one A
//             and one B fragment from LDS.
// At the end of a row the tiles are written to LDS as Ypre, folded (a CSR
// table of (h, t) pairs per packed output) and written once to Y.
#include "sp_internal.hpp"

#include <algorithm>

namespace spamd {

namespace {

constexpr int cbinom(int n, int k) {
  if (k < 0 || k > n) return 0;
  long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return (int)r;
}
constexpr int psz(int R, int k) { return cbinom(R + k - 1, k); }

constexpr int pidx(const int *t, int k, int R) {
  int off = 0, start = 0;
  for (int j = 0; j < k; ++j) {
    for (int i = start; i < t[j]; ++i) off += psz(R - i, k - 1 - j);
    start = t[j];
  }
  return off;
}

// Gather table of the product of a degree-K polynomial with a linear form:
// output J (degree K+1) = sum over its distinct values j of p[J - e_j] u[j].
template <int R, int K>
struct ProdTab {
  static constexpr int PO = psz(R, K + 1);
  int n[PO];
  int src[PO][K + 1];
  int jj[PO][K + 1];
  constexpr ProdTab() : n{}, src{}, jj{} {
    int idx[K + 1] = {};
    for (int p = 0; p < PO; ++p) {
      int cnt = 0;
      for (int q = 0; q <= K; ++q) {
        if (q > 0 && idx[q] == idx[q - 1]) continue;
        int sub[K + 1] = {};
        int w = 0;
        for (int r = 0; r <= K; ++r)
          if (r != q) sub[w++] = idx[r];
        src[p][cnt] = pidx(sub, K, R);
        jj[p][cnt] = idx[q];
        ++cnt;
      }
      n[p] = cnt;
      int pos = K;
      while (pos >= 0 && idx[pos] == R - 1) --pos;
      if (pos < 0) break;
      idx[pos]++;
      for (int q = pos + 1; q <= K; ++q) idx[q] = idx[pos];
    }
  }
};

template <int R, int K, class T>
__device__ __forceinline__ T prod_entry(const T *p, const T *u, int J) {
  constexpr ProdTab<R, K> tab{};
  T s = 0;
#pragma unroll
  for (int q = 0; q < K + 1; ++q)
    if (q < tab.n[J]) s = fma(p[tab.src[J][q]], u[tab.jj[J][q]], s);
  return s;
}

// All coefficients of degree K+1 into registers.
template <int R, int K, class T>
__device__ __forceinline__ void prod_full(const T *p, const T *u, T *out) {
#pragma unroll
  for (int J = 0; J < psz(R, K + 1); ++J) out[J] = prod_entry<R, K>(p, u, J);
}

// Quarter W of the degree-D polynomial of rows rows[0..D-1] (each R values),
// scaled: store(J, value) for J in [lo, hi) of the quarter, and store(J, 0)
// for the padding entries [P, padded) (handled by the last quarter).
template <int R, int D, int W, class T, class Store>
__device__ __forceinline__ void poly_quarter(const T *rows, T scale, Store &&store,
                                             int padded) {
  constexpr int P = psz(R, D);
  constexpr int Q = (P + 3) / 4;
  constexpr int lo = W * Q;
  constexpr int hi = (W + 1) * Q < P ? (W + 1) * Q : P;
  if constexpr (D == 1) {
#pragma unroll
    for (int J = lo; J < hi; ++J) store(J, rows[J] * scale);
  } else if constexpr (D == 2) {
#pragma unroll
    for (int J = lo; J < hi; ++J) store(J, prod_entry<R, 1>(rows, rows + R, J) * scale);
  } else {
    static_assert(D == 3, "head/tail degree above 3 is not instantiated");
    T p2[psz(R, 2)];
    prod_full<R, 1>(rows, rows + R, p2);
#pragma unroll
    for (int J = lo; J < hi; ++J) store(J, prod_entry<R, 2>(p2, rows + 2 * R, J) * scale);
  }
  if constexpr (W == 3) {
    for (int J = P; J < padded; ++J) store(J, T(0));
  }
}

// Quarter `wave` of the head (scaled by c) and of the tail (scaled by one).
template <int R, int H, int TD, class T, class SA, class SB>
__device__ __forceinline__ void produce_quarters(int wave, const T *rows, T c, T one, SA &&sa,
                                                 SB &&sb, int padA, int padB) {
  switch (wave) {
    case 0:
      poly_quarter<R, H, 0>(rows, c, sa, padA);
      poly_quarter<R, TD, 0>(rows + H * R, one, sb, padB);
      break;
    case 1:
      poly_quarter<R, H, 1>(rows, c, sa, padA);
      poly_quarter<R, TD, 1>(rows + H * R, one, sb, padB);
      break;
    case 2:
      poly_quarter<R, H, 2>(rows, c, sa, padA);
      poly_quarter<R, TD, 2>(rows + H * R, one, sb, padB);
      break;
    default:
      poly_quarter<R, H, 3>(rows, c, sa, padA);
      poly_quarter<R, TD, 3>(rows + H * R, one, sb, padB);
      break;
  }
}

struct FastArgs {
  const Chunk *chunks;
  const Batch *batches;
  const u32 *tup;
  const double *coef;
  const double *U;
  const float *coef32;
  const float *U32;
  const u32 *fold_ptr;   // P_M + 1: CSR fold lists
  const u32 *fold_term;  // Ypre offsets h * P_T + t
  const u32 *fold_tab;   // [fold_len][256]: the same terms, thread-major
  int fold_len;
  const double *fact;    // prod f_J! of the P_M outputs
  double *Y;
  double *partial;
};

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Element type of the fast kernel: FP64 operands on v_mfma_f64_16x16x4f64,
// or FP32 operands on v_mfma_f32_16x16x4f32.  Both take lane l's A[l%16][l/16]
// and B[l/16][l%16]; accumulator register q holds C[4q + l/16][l%16] for f64
// and C[4(l/16) + q][l%16] for f32.
template <class T>
struct FastElem;
template <>
struct FastElem<double> {
  using acc_t = d4;
  static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int crow(int q, int kr) { return 4 * q + kr; }
  static __device__ __forceinline__ const double *coef(const struct FastArgs &a);
  static __device__ __forceinline__ const double *U(const struct FastArgs &a);
};
template <>
struct FastElem<float> {
  using acc_t = f4;
  static __device__ __forceinline__ acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int crow(int q, int kr) { return 4 * kr + q; }
  static __device__ __forceinline__ const float *coef(const struct FastArgs &a);
  static __device__ __forceinline__ const float *U(const struct FastArgs &a);
};
__device__ __forceinline__ const double *FastElem<double>::coef(const FastArgs &a) { return a.coef; }
__device__ __forceinline__ const double *FastElem<double>::U(const FastArgs &a) { return a.U; }
__device__ __forceinline__ const float *FastElem<float>::coef(const FastArgs &a) { return a.coef32; }
__device__ __forceinline__ const float *FastElem<float>::U(const FastArgs &a) { return a.U32; }

struct BatchV {
  u64 o0;
  u32 row;
  int nb;
  bool last;
};

// Uniform batch descriptor.  readfirstlane keeps every field in scalar
// registers at the point of the load, so the load is not sunk into the
// divergent owner branch (where it would become a vector load whose wait
// also drains the row prefetch).
__device__ __forceinline__ BatchV read_batch(const Batch *__restrict__ bs, u64 b) {
  const Batch x = bs[b];
  const u32 lo = __builtin_amdgcn_readfirstlane(x.o0_lo);
  const u32 hi = __builtin_amdgcn_readfirstlane(x.o0_hi);
  const u32 row = __builtin_amdgcn_readfirstlane(x.row);
  const u32 nl = __builtin_amdgcn_readfirstlane(x.nb_last);
  return BatchV{((u64)hi << 32) | lo, row, (int)(nl & 0xFFFFu), (nl >> 31) != 0};
}

// Loads the owner of this lane in batch `bt` (tuple and coefficient) without
// a branch, so the loads stay in flight until first use: lanes past the
// batch read a valid owner (the batch's first, or owner 0 for an empty
// batch) and get coefficient 0.
template <int M, class T>
__device__ __forceinline__ void load_owner(const FastArgs &a, const T *coef, const BatchV &bt,
                                           int lane, u32 (&tp)[M], T &cv, bool &in) {
  in = lane < bt.nb;
  const u64 o = in ? bt.o0 + lane : (bt.nb > 0 ? bt.o0 : 0);
  const u32 *t = a.tup + o * M;
#pragma unroll
  for (int k = 0; k < M; ++k) tp[k] = t[k];
  cv = coef[o];
}

template <int M, int R, class T>
__device__ __forceinline__ void load_rows(const T *U, const u32 (&tp)[M], T (&rows)[M * R]) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const T *ur = U + (u64)tp[k] * R;
#pragma unroll
    for (int j = 0; j < R; ++j) rows[k * R + j] = ur[j];
  }
}

// Fold of a row's Ypre (in LDS) into its packed outputs: output J sums the
// Ypre entries (h, t) whose merged multi-index is J (the sum runs in (h, t)
// order), then takes the factor prod(f_J!).  Two forms of the same
// arithmetic:
//  - CSR (fold_ptr / fold_term): thread x takes outputs x, x + 256, ... and
//    walks each one's term list (dependent loads: pointer, then terms);
//  - thread-major (fold_tab, fold_len): thread x reads its terms for all its
//    outputs from fold_tab[i * 256 + x], i < fold_len (output slot k =
//    J / 256 in bits 28..31, the Ypre offset h * P_T + t below, slot 15
//    pads), coalesced loads that do not depend on one another.
// Measured on MI355X: the thread-major form makes the BF16 kernel 4-11%
// faster on orders 4-7, the order-7 FP64 kernel (C(6, 3) = 20 terms per
// output) 4% and the order-6 FP64 kernel (C(5, 3) = 10 terms, with the tail
// strip) 2% faster; with shorter lists it costs the FP64 kernels registers
// (the order-5 FP64 kernel went from 128 to 152 VGPRs and 9% slower) and the
// FP32 kernels gain nothing, so those keep the CSR form.
template <int M, int H, class TE>
constexpr bool fold_thread_major() {
  return sizeof(TE) == 2 || (sizeof(TE) == 8 && cbinom(M, H) >= 10);
}

template <int M, int H, int PM, class TE, class TF>
__device__ __forceinline__ void fold_row(const FastArgs &a, const TF *sF, double *out) {
  if constexpr (fold_thread_major<M, H, TE>()) {
    constexpr int KM = (PM + 255) / 256;
    static_assert(KM < 15, "fold slots");
    double s[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) s[k] = 0.0;
    const u32 *ft = a.fold_tab + threadIdx.x;
#pragma unroll 4
    for (int i = 0; i < a.fold_len; ++i) {
      const u32 e = ft[(u64)i * 256];
      const double v = (double)sF[e & 0x0FFFFFFFu];
      const int k = (int)(e >> 28);
#pragma unroll
      for (int kk = 0; kk < KM; ++kk)
        if (kk == k) s[kk] += v;
    }
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int J = threadIdx.x + 256 * k;
      if (J < PM) out[J] = s[k] * a.fact[J];
    }
  } else {
    for (int J = threadIdx.x; J < PM; J += 256) {
      double s = 0.0;
      for (u32 q = a.fold_ptr[J]; q < a.fold_ptr[J + 1]; ++q)
        s += (double)sF[a.fold_term[q]];
      out[J] = s * a.fact[J];
    }
  }
}

// Owners per batch of the FP64 kernel: 64, or 48 where the FP64
// operand tiles of 64 owners would exceed 80 KB of LDS, so that two
// workgroups fit a CU (160 KB) and one's producer phase overlaps the other's
// MFMA phase.
constexpr int fast_bo(int M, int R, int H) {
  return (64 * ((psz(R, H) | 1) + (psz(R, M - H) | 1)) + 16) * 8 > 80 * 1024 ? 48 : 64;
}

// TE = double: the FP64 form; TE = float: the FP32 mixed form (FP32 U rows,
// coefficients, products and accumulation; the fold sums in FP64).
template <int M, int R, int H, class TE>
constexpr int kernel_bo() {
  return sizeof(TE) == 8 ? fast_bo(M, R, H) : 64;  // the FP32 form runs 64-owner batches
}

// Tile ownership of the four waves.  When the TH x TT tile grid splits into
// four equal rectangles (sh x st = 4 with sh | TH, st | TT), wave w owns the
// rectangle (w / st, w % st) and per k-step reads each of its TH/sh A and
// TT/st B fragments once for all (TH/sh)*(TT/st) MFMAs; the split with the
// fewest fragment reads is taken.  Otherwise tiles go round-robin (tile
// w + 4i) with one A and one B read per MFMA.  Measured on MI355X: the 2 x 2
// split makes the order-7 rank-6 FP64 kernel 9% faster and both splits make
// the FP32 kernels 10-12% faster, but the 4 x 1 strips leave the order-6
// rank-8 FP64 kernel 0.5% slower, so FP64 kernels (strips = false) take
// only the 2 x 2 split.
struct TileSplit {
  int sh, st;  // 0, 0: round-robin
};
constexpr TileSplit tile_split(int TH, int TT, bool strips) {
  TileSplit best{0, 0};
  int reads = 1 << 30;
  const int opts[3][2] = {{4, 1}, {2, 2}, {1, 4}};
  for (const auto &o : opts) {
    if (TH % o[0] != 0 || TT % o[1] != 0) continue;
    if (!strips && o[0] != 2) continue;
    const int r = TH / o[0] + TT / o[1];
    if ((TH / o[0]) * (TT / o[1]) > 1 && r < reads) {
      reads = r;
      best = TileSplit{o[0], o[1]};
    }
  }
  return best;
}

template <int M, int R, int H, class TE>
__global__ void __launch_bounds__(256, (kernel_bo<M, R, H, TE>() < 64 ? 2 : 1))
s3ttmc_mfma(FastArgs a) {
  using EL = FastElem<TE>;
  using acc_t = typename EL::acc_t;
  constexpr int T = M - H;
  constexpr int PH = psz(R, H), PT = psz(R, T), PM = psz(R, M);
  // Tail strip: when the tail's last 16-wide tile would hold only 1-4 real
  // columns next to two full ones (P_T = 33..36: orders 5 and 6 at R = 8),
  // the MFMAs cover the first 32 columns and those NSC columns are summed on
  // the VALU from the same A fragments (the vector and matrix pipes run side
  // by side): a third fewer MFMAs for a few FMAs per fragment.
  constexpr bool STRIP = PT > 32 && PT <= 36;
  constexpr int NSC = STRIP ? PT - 32 : 0;
  constexpr int TH = (PH + 15) / 16, TT = STRIP ? 2 : (PT + 15) / 16;
  constexpr int NTILE = TH * TT;
  constexpr int TPW = (NTILE + 3) / 4;
  constexpr int BO = kernel_bo<M, R, H, TE>();
  // Operand rows hold exactly P_H (P_T) coefficients, rounded up to an odd
  // stride that spreads the 4 row groups over the LDS banks.  A fragment read
  // in the padding of the last 16-wide tile returns a neighbouring owner's
  // value; it only reaches accumulator rows h >= P_H (columns t >= P_T),
  // which the fold never reads.  TAIL elements after the tiles keep the last
  // owner's padding reads inside the array.
  constexpr int LDA = PH | 1;
  constexpr int LDB = PT | 1;
  constexpr int TAIL = 16;
  constexpr int SA = BO * LDA, SB = BO * LDB;
  constexpr int SF = PH * PT;
  constexpr int SMEM = (SA + SB + TAIL) > SF ? (SA + SB + TAIL) : SF;
  __shared__ TE sm[SMEM];
  TE *sA = sm;
  TE *sB = sm + SA;
  TE *sF = sm;

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kr = lane >> 4, kc = lane & 15;
  const Chunk ch = a.chunks[blockIdx.x];
  if (ch.b_begin >= ch.b_end) return;

  constexpr TileSplit SPL = tile_split(TH, TT, sizeof(TE) == 4);
  constexpr bool BLOCKED = SPL.sh > 0;
  constexpr int BH = BLOCKED ? TH / SPL.sh : 1;  // row tiles of a wave's rectangle
  constexpr int BT = BLOCKED ? TT / SPL.st : 1;  // column tiles
  static_assert(!BLOCKED || BH * BT == TPW, "tile rectangle");
  // the tile of accumulator i: (row tile, column tile)
  auto tile_of = [&](int i, int &th, int &tt) {
    if constexpr (BLOCKED) {
      th = (wave / SPL.st) * BH + i / BT;
      tt = (wave % SPL.st) * BT + i % BT;
    } else {
      const int tile = wave + 4 * i;
      th = tile / TT;
      tt = tile % TT;
    }
  };
  // LDS offsets of this lane's A and B fragments for the wave's tiles
  int aoff[TPW], boff[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    int th, tt;
    tile_of(i, th, tt);
    if (!BLOCKED && wave + 4 * i >= NTILE) th = (NTILE - 1) / TT, tt = (NTILE - 1) % TT;
    aoff[i] = kr * LDA + th * 16 + kc;
    boff[i] = kr * LDB + tt * 16 + kc;
  }

  // Software pipeline (loop-carried registers):
  //   rows, cur_c      U rows and coefficient of batch b
  //   tpn, nxt_cv/in   tuple and raw coefficient of batch b+1
  // The loads for b+1 (rows) and b+2 (tuple) are issued after the first
  // barrier of iteration b and first used in iteration b+1, so they are in
  // flight during the MFMA phase.  Batch indices past the chunk are clamped
  // to its last batch (valid memory, results unused).
  const u64 blast = ch.b_end - 1;
  u32 tpn[M];
  TE cur_c, nxt_cv;
  bool nxt_in;
  TE rows[M * R];
  const TE *coef = EL::coef(a);
  const TE *Urows = EL::U(a);
  {
    u32 tp0[M];
    TE cv0;
    bool in0;
    load_owner<M>(a, coef, read_batch(a.batches, ch.b_begin), lane, tp0, cv0, in0);
    cur_c = in0 ? cv0 : TE(0);
    load_rows<M, R>(Urows, tp0, rows);
    load_owner<M>(a, coef, read_batch(a.batches, min(ch.b_begin + 1, blast)), lane, tpn,
                  nxt_cv, nxt_in);
  }

  acc_t acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) acc[i] = acc_t{0, 0, 0, 0};
  // Strip: with TT = 2, round-robin tiles or 2 x 2 rectangles give the waves
  // of a pair (w, w^1) the A fragments of the same head tiles (tile column
  // w % 2), and wave w sums half the strip columns, c0 .. c0 + NSCW - 1, for
  // each of them; with 4 x 1 rectangles every wave holds its own head tiles
  // and sums all strip columns.  Lane (kr, kc) accumulates head coefficient
  // h = 16 th + kc over the owners 4 ks + kr; the four partial sums of an h
  // are added across kr at the end of the row.
  constexpr bool PAIR = !BLOCKED || SPL.st == 2;
  constexpr int NSCW = PAIR ? (NSC + 1) / 2 : NSC;  // strip columns of a wave
  constexpr int NSL = BLOCKED ? BH : TPW;  // head tiles of a wave
  const int c0 = PAIR ? (wave & 1) * NSCW : 0;
  TE strip[STRIP ? NSL : 1][STRIP ? NSCW : 1];
  if constexpr (STRIP) {
#pragma unroll
    for (int u = 0; u < NSL; ++u)
#pragma unroll
      for (int j = 0; j < NSCW; ++j) strip[u][j] = TE(0);
  }

  for (u64 b = ch.b_begin; b < ch.b_end; ++b) {
    const BatchV bt = read_batch(a.batches, b);
    const int nb = bt.nb;
    const int ksteps = (nb + 3) >> 2;
    // ---- producer: lane = owner, wave = quarter of the coefficients
    if (lane < ksteps * 4) {
      TE *arow = sA + lane * LDA;
      TE *brow = sB + lane * LDB;
      // padding lanes (lane >= nb) write zero rows: both scales are 0
      const TE one = lane < nb ? TE(1) : TE(0);
      produce_quarters<R, H, T>(
          wave, rows, cur_c, one, [&](int J, TE v) { arow[J] = v; },
          [&](int J, TE v) { brow[J] = v; }, PH, PT);
    }
    __syncthreads();
    // ---- prefetch: U rows of batch b+1, tuple of batch b+2
    cur_c = nxt_in ? nxt_cv : TE(0);
    load_rows<M, R>(Urows, tpn, rows);
    load_owner<M>(a, coef, read_batch(a.batches, min(b + 2, blast)), lane, tpn, nxt_cv,
                  nxt_in);
    // ---- consumer: MFMA over the batch
    for (int ks = 0; ks < ksteps; ++ks) {
      const TE *ar = sA + ks * 4 * LDA;
      const TE *br = sB + ks * 4 * LDB;
      // this lane's owner's strip columns (zero past the strip)
      TE bs[STRIP ? NSCW : 1];
      if constexpr (STRIP) {
#pragma unroll
        for (int j = 0; j < NSCW; ++j)
          bs[j] = c0 + j < NSC ? br[kr * LDB + 32 + c0 + j] : TE(0);
      }
      if constexpr (BLOCKED) {
        // each fragment read once per k-step (accumulator i = u * BT + v)
        TE af[BH], bf[BT];
#pragma unroll
        for (int u = 0; u < BH; ++u) af[u] = ar[aoff[u * BT]];
#pragma unroll
        for (int v = 0; v < BT; ++v) bf[v] = br[boff[v]];
#pragma unroll
        for (int u = 0; u < BH; ++u)
#pragma unroll
          for (int v = 0; v < BT; ++v) acc[u * BT + v] = EL::mma(af[u], bf[v], acc[u * BT + v]);
        if constexpr (STRIP) {
#pragma unroll
          for (int u = 0; u < BH; ++u)
#pragma unroll
            for (int j = 0; j < NSCW; ++j) strip[u][j] = fma(af[u], bs[j], strip[u][j]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
          if (NTILE % 4 == 0 || wave + 4 * i < NTILE) {
            const TE av = ar[aoff[i]];
            acc[i] = EL::mma(av, br[boff[i]], acc[i]);
            if constexpr (STRIP) {
#pragma unroll
              for (int j = 0; j < NSCW; ++j) strip[i][j] = fma(av, bs[j], strip[i][j]);
            }
          }
        }
      }
    }
    __syncthreads();
    if (!bt.last) continue;
    // ---- fold Ypre into the packed row
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      if (BLOCKED || wave + 4 * i < NTILE) {
        int th, tt;
        tile_of(i, th, tt);
        const int t = tt * 16 + kc;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int h = th * 16 + EL::crow(q, kr);
          if (h < PH && t < PT) sF[h * PT + t] = acc[i][q];
        }
      }
      acc[i] = acc_t{0, 0, 0, 0};
    }
    if constexpr (STRIP) {
      // strip column c0 + j of head h = 16 th + kc: the four owner classes
      // kr added in a fixed order ((0 + 1) + (2 + 3) in every lane)
#pragma unroll
      for (int u = 0; u < NSL; ++u) {
        int th, tt;
        tile_of(BLOCKED ? u * BT : u, th, tt);
        const bool valid = BLOCKED || wave + 4 * u < NTILE;
        const int h = th * 16 + kc;
#pragma unroll
        for (int j = 0; j < NSCW; ++j) {
          double v = strip[u][j];
          v += __shfl_xor(v, 16);
          v += __shfl_xor(v, 32);
          if (valid && kr == 0 && h < PH && c0 + j < NSC) sF[h * PT + 32 + c0 + j] = v;
          strip[u][j] = TE(0);
        }
      }
    }
    __syncthreads();
    double *out = (ch.partial >= 0) ? a.partial + (u64)ch.partial * PM : a.Y + (u64)bt.row * PM;
    fold_row<M, H, PM, TE>(a, sF, out);
    __syncthreads();
  }
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Mixed-precision form of s3ttmc_mfma: FP32 U rows, coefficients and
// polynomial products; A = c*head and B = tail rounded to BF16 (round to
// nearest even) and multiplied on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation; the fold sums in FP64.  Operands are stored transposed
// (coefficient-major, owner-minor) so that one 16-byte LDS read gives a
// lane its 8 consecutive owners of the K dimension.  A batch of 64 owners is
// two K-steps of 32.
template <int M, int R, int H>
__global__ void __launch_bounds__(256)
s3ttmc_mfma_bf16(FastArgs a) {
  constexpr int T = M - H;
  constexpr int PH = psz(R, H), PT = psz(R, T), PM = psz(R, M);
  constexpr int TH = (PH + 15) / 16, TT = (PT + 15) / 16;
  constexpr int NTILE = TH * TT;
  constexpr int TPW = (NTILE + 3) / 4;
  constexpr int BO = 64;
  constexpr int LDT = BO + 8;  // bf16 elements per coefficient row (16 B pad)
  __shared__ __attribute__((aligned(16))) __bf16 sA[TH * 16 * LDT];
  __shared__ __attribute__((aligned(16))) __bf16 sB[TT * 16 * LDT];
  __shared__ float sF[PH * PT];

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kr = lane >> 4, kc = lane & 15;
  const Chunk ch = a.chunks[blockIdx.x];
  if (ch.b_begin >= ch.b_end) return;

  int aoff[TPW], boff[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int tile = wave + 4 * i < NTILE ? wave + 4 * i : NTILE - 1;
    aoff[i] = ((tile / TT) * 16 + kc) * LDT + 8 * kr;
    boff[i] = ((tile % TT) * 16 + kc) * LDT + 8 * kr;
  }

  const u64 blast = ch.b_end - 1;
  u32 tpn[M];
  float cur_c, nxt_cv;
  bool nxt_in;
  float rows[M * R];
  {
    u32 tp0[M];
    float cv0;
    bool in0;
    load_owner<M>(a, a.coef32, read_batch(a.batches, ch.b_begin), lane, tp0, cv0, in0);
    cur_c = in0 ? cv0 : 0.f;
    load_rows<M, R>(a.U32, tp0, rows);
    load_owner<M>(a, a.coef32, read_batch(a.batches, min(ch.b_begin + 1, blast)), lane, tpn,
                  nxt_cv, nxt_in);
  }

  f4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};

  for (u64 b = ch.b_begin; b < ch.b_end; ++b) {
    const BatchV bt = read_batch(a.batches, b);
    const int nb = bt.nb;
    const int ksteps = (nb + 31) >> 5;
    // ---- producer: every lane writes its owner column (zeros past nb)
    if (nb > 0) {
      const float one = lane < nb ? 1.f : 0.f;
      produce_quarters<R, H, T>(
          wave, rows, cur_c, one, [&](int J, float v) { sA[J * LDT + lane] = (__bf16)v; },
          [&](int J, float v) { sB[J * LDT + lane] = (__bf16)v; }, TH * 16, TT * 16);
    }
    __syncthreads();
    cur_c = nxt_in ? nxt_cv : 0.f;
    load_rows<M, R>(a.U32, tpn, rows);
    load_owner<M>(a, a.coef32, read_batch(a.batches, min(b + 2, blast)), lane, tpn, nxt_cv,
                  nxt_in);
    for (int ks = 0; ks < ksteps; ++ks) {
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        if (NTILE % 4 == 0 || wave + 4 * i < NTILE) {
          const bf8 av = *reinterpret_cast<const bf8 *>(sA + aoff[i] + ks * 32);
          const bf8 bv = *reinterpret_cast<const bf8 *>(sB + boff[i] + ks * 32);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    if (!bt.last) continue;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int tile = wave + 4 * i;
      if (tile < NTILE) {
        const int th = tile / TT, tt = tile % TT;
        const int t = tt * 16 + kc;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int h = th * 16 + 4 * kr + q;
          if (h < PH && t < PT) sF[h * PT + t] = acc[i][q];
        }
      }
      acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();
    double *out = (ch.partial >= 0) ? a.partial + (u64)ch.partial * PM : a.Y + (u64)bt.row * PM;
    fold_row<M, H, PM, __bf16>(a, sF, out);
    __syncthreads();
  }
}

using FastKernel = void (*)(FastArgs);

struct FastEntry {
  int M, R, H;
  int bo;  // owners per batch
  FastKernel k64, kf32, kbf16;
};

#define SP_FAST(M, R, H)                                                        \
  FastEntry {                                                                   \
    M, R, H, fast_bo(M, R, H), s3ttmc_mfma<M, R, H, double>,                     \
        s3ttmc_mfma<M, R, H, float>, s3ttmc_mfma_bf16<M, R, H>                  \
  }

const FastEntry kFast[] = {
    // order 3, R = 2..16
    SP_FAST(2, 2, 1), SP_FAST(2, 3, 1), SP_FAST(2, 4, 1), SP_FAST(2, 5, 1), SP_FAST(2, 6, 1),
    SP_FAST(2, 7, 1), SP_FAST(2, 8, 1), SP_FAST(2, 9, 1), SP_FAST(2, 10, 1), SP_FAST(2, 11, 1),
    SP_FAST(2, 12, 1), SP_FAST(2, 13, 1), SP_FAST(2, 14, 1), SP_FAST(2, 15, 1), SP_FAST(2, 16, 1),
    // order 4, R = 2..16
    SP_FAST(3, 2, 2), SP_FAST(3, 3, 2), SP_FAST(3, 4, 2), SP_FAST(3, 5, 2), SP_FAST(3, 6, 2),
    SP_FAST(3, 7, 2), SP_FAST(3, 8, 2), SP_FAST(3, 9, 2), SP_FAST(3, 10, 2), SP_FAST(3, 11, 2),
    SP_FAST(3, 12, 2), SP_FAST(3, 13, 2), SP_FAST(3, 14, 2), SP_FAST(3, 15, 2), SP_FAST(3, 16, 2),
    // order 5, R = 2..12
    SP_FAST(4, 2, 2), SP_FAST(4, 3, 2), SP_FAST(4, 4, 2), SP_FAST(4, 5, 2), SP_FAST(4, 6, 2),
    SP_FAST(4, 7, 2), SP_FAST(4, 8, 2), SP_FAST(4, 9, 2), SP_FAST(4, 10, 2), SP_FAST(4, 11, 2),
    SP_FAST(4, 12, 2),
    // order 6, R = 2..8
    SP_FAST(5, 2, 3), SP_FAST(5, 3, 3), SP_FAST(5, 4, 3), SP_FAST(5, 5, 3), SP_FAST(5, 6, 3),
    SP_FAST(5, 7, 3), SP_FAST(5, 8, 3),
    // order 7, R = 2..6
    SP_FAST(6, 2, 3), SP_FAST(6, 3, 3), SP_FAST(6, 4, 3), SP_FAST(6, 5, 3), SP_FAST(6, 6, 3),
};

}  // namespace

// Fold table of (h, t) pairs per packed output of degree M for head degree H.
void make_fold_table(u32 R, int M, int H, std::vector<u32> &ptr, std::vector<u32> &term,
                     std::vector<u32> &tab, int &len) {
  const int T = M - H;
  const PackedTables th = make_tables(R, H), tt = make_tables(R, T), tm = make_tables(R, M);
  auto enumerate = [&](int k) {
    std::vector<std::vector<u32>> out;
    std::vector<u32> idx(k, 0);
    while (true) {
      out.push_back(idx);
      int p = k - 1;
      while (p >= 0 && idx[p] == R - 1) --p;
      if (p < 0) break;
      idx[p]++;
      for (int q = p + 1; q < k; ++q) idx[q] = idx[p];
    }
    return out;
  };
  const auto eh = enumerate(H), et = enumerate(T);
  auto index_of = [&](const std::vector<u32> &t) {
    u64 off = 0;
    u32 start = 0;
    const int k = (int)t.size();
    for (int j = 0; j < k; ++j) {
      for (u32 i = start; i < t[j]; ++i) off += symsize(R - i, k - 1 - j);
      start = t[j];
    }
    return off;
  };
  std::vector<std::vector<u32>> lists(tm.P);
  for (u64 h = 0; h < th.P; ++h)
    for (u64 t = 0; t < tt.P; ++t) {
      std::vector<u32> J(eh[h]);
      J.insert(J.end(), et[t].begin(), et[t].end());
      std::sort(J.begin(), J.end());
      lists[index_of(J)].push_back((u32)(h * tt.P + t));
    }
  ptr.assign(1, 0);
  term.clear();
  for (const auto &l : lists) {
    term.insert(term.end(), l.begin(), l.end());
    ptr.push_back((u32)term.size());
  }
  // thread x: the terms of outputs x, x + 256, ... (slot k = J / 256), each
  // output's terms in (h, t) order; lists padded to the longest with slot 15
  constexpr u32 kPad = 15u << 28;
  if (tm.P > 14 * 256 || th.P * tt.P >= (1u << 28))
    throw Error(SP_ERR_UNSUPPORTED, "fold table too large");
  std::vector<std::vector<u32>> per(256);
  for (u64 J = 0; J < tm.P; ++J)
    for (u32 off : lists[J]) per[J % 256].push_back((u32)((J / 256) << 28) | off);
  len = 0;
  for (const auto &l : per) len = std::max(len, (int)l.size());
  tab.assign((size_t)std::max(len, 1) * 256, kPad);
  for (int x = 0; x < 256; ++x)
    for (size_t i = 0; i < per[x].size(); ++i) tab[i * 256 + x] = per[x][i];
}

// Returns the instantiated FP64 fast kernel for (M, R) or null; sets H and
// the FP32 and BF16 forms of the same shape.
void *fast_kernel_for(int M, u32 R, int *H, int *bo, void **kf32, void **kbf16) {
  for (const auto &e : kFast)
    if (e.M == M && (u32)e.R == R) {
      *H = e.H;
      if (bo) *bo = e.bo;
      if (kf32) *kf32 = (void *)e.kf32;
      if (kbf16) *kbf16 = (void *)e.kbf16;
      return (void *)e.k64;
    }
  return nullptr;
}

void launch_fast(void *kernel, unsigned grid, hipStream_t s, const Chunk *chunks,
                 const Batch *batches, const u32 *tup, const double *coef, const double *U,
                 const float *coef32, const float *U32, const u32 *fold_ptr,
                 const u32 *fold_term, const u32 *fold_tab, int fold_len, const double *fact,
                 double *Y, double *partial) {
  FastArgs a{chunks,    batches,  tup,      coef,     U,    coef32, U32,
             fold_ptr,  fold_term, fold_tab, fold_len, fact, Y,      partial};
  ((FastKernel)kernel)<<<grid, 256, 0, s>>>(a);
}

}  // namespace spamd
