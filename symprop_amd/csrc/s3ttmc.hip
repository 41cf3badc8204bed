// S3TTMc on MI355X: Y = X x_2 U ... x_N U, packed along the last N-1 modes.
//
// Reference semantics (src/compressed_dtree.cpp:421-518): for every owner
// (leaf L = M \ {a_j}, rem = a_j, value v) of every nonzero M,
//     Y[rem, J] += v * N!/prod(f_M!) / N! * K(L)[J],   J a sorted (N-1)-tuple,
// where K(L) is the leaf tensor that the reference propagates up the tree by
// exclude-one one-sided outer products (symtensor.hpp:24-39): K(L)[J] is the
// permanent of U[L, J].
//
// Formulation used here.  The permanent of U[L, J] over a sorted multi-index
// J with multiplicities f_J equals prod(f_J!) times the coefficient of the
// monomial x^J in the polynomial  prod_{a in L} (u_a . x),  u_a = U row of
// index a.  So K(L) follows from the prefix chain p_1 = u_{l_0},
// p_k = p_{k-1} * (u_{l_{k-1}} . x) alone, without the exclude-one parents and
// without materialising tree levels.  In packed lexicographic order the
// product is a gather: p_k[J] = sum over the distinct values j in J of
// p_{k-1}[J - e_j] * u[j].
//
// Reduction.  Owners are sorted by rem (then leaf); a work chunk covers whole
// Y rows, or one piece of a row with many owners.  A chunk accumulates its
// rows in on-chip memory and writes each packed Y row exactly once (split
// rows are summed from their pieces in a fixed order), so there are no
// atomics and repeated calls are bit-identical.
#include "sp_internal.hpp"

#include <algorithm>
#include <cmath>

namespace spamd {

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

u64 symsize(u64 R, u64 k) { return binom(k + R - 1, k); }

namespace {

u64 packed_index(const std::vector<u32> &t, u32 R) {
  const int k = (int)t.size();
  u64 off = 0;
  u32 start = 0;
  for (int j = 0; j < k; ++j) {
    for (u32 i = start; i < t[j]; ++i) off += symsize(R - i, k - 1 - j);
    start = t[j];
  }
  return off;
}

double factorial_d(int n) {
  double f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

}  // namespace

PackedTables make_tables(u32 R, int k) {
  PackedTables T;
  T.R = R;
  T.k = k;
  T.P = symsize(R, k);
  T.term_ptr.push_back(0);
  std::vector<u32> idx(k, 0);
  while (true) {
    // gather terms: distinct values j of the tuple, src = tuple minus one j
    if (k >= 2) {
      for (int q = 0; q < k; ++q) {
        if (q > 0 && idx[q] == idx[q - 1]) continue;
        std::vector<u32> sub;
        for (int r = 0; r < k; ++r)
          if (r != q) sub.push_back(idx[r]);
        T.term.push_back((u32)(packed_index(sub, R) << 8) | idx[q]);
      }
      T.term_ptr.push_back((u32)T.term.size());
    }
    double fp = 1.0;
    int run = 1;
    for (int q = 1; q <= k; ++q) {
      if (q < k && idx[q] == idx[q - 1]) ++run;
      else { fp *= factorial_d(run); run = 1; }
    }
    T.fact_prod.push_back(fp);
    T.weight.push_back(factorial_d(k) / fp);
    int p = k - 1;
    while (p >= 0 && idx[p] == R - 1) --p;
    if (p < 0) break;
    idx[p]++;
    for (int q = p + 1; q < k; ++q) idx[q] = idx[p];
  }
  return T;
}

namespace {

constexpr int GEN_BLOCK = 64;

struct GenArgs {
  const Chunk *chunks;
  const u32 *tup;
  const u64 *segptr;
  const u32 *tab_ptr;    // concatenated per level k = 2..m
  const u32 *tab_term;
  const double *fact;    // prod f! for the P_m outputs
  u64 ptr_off[17];       // offsets into tab_ptr per k
  u64 term_off[17];
  u64 chain_off[17];     // LDS offsets (elements) of p_k, k = 1..m-1
  u64 Pk[17];
  u64 acc_off;
  u64 u_off;
  int m;
  u32 R;
  u64 P;
  double *Y;
  double *partial;
};

// Generic kernel: one wave per chunk, any order/rank that fits in LDS.
template <class T>
__global__ void __launch_bounds__(GEN_BLOCK)
s3ttmc_generic(GenArgs a, const T *__restrict__ U, const T *__restrict__ coef) {
  extern __shared__ unsigned char smem_raw[];
  T *sm = reinterpret_cast<T *>(smem_raw);
  const Chunk ch = a.chunks[blockIdx.x];
  const int m = a.m;
  const u32 R = a.R;
  T *acc = sm + a.acc_off;
  T *urow = sm + a.u_off;
  for (u32 r = ch.r_begin; r < ch.r_end; ++r) {
    u64 ob, oe;
    if (ch.partial >= 0) { ob = ch.o_begin; oe = ch.o_end; }
    else { ob = a.segptr[r]; oe = a.segptr[r + 1]; }
    for (u64 J = threadIdx.x; J < a.P; J += GEN_BLOCK) acc[J] = 0;
    for (u64 o = ob; o < oe; ++o) {
      __syncthreads();
      for (u32 t = threadIdx.x; t < (u32)m * R; t += GEN_BLOCK) {
        const u32 row = a.tup[o * m + t / R];
        urow[t] = U[(u64)row * R + t % R];
      }
      __syncthreads();
      const T *prev = urow;  // p_1 = u_{l_0}
      for (int k = 2; k < m; ++k) {
        T *dst = sm + a.chain_off[k];
        const u32 *tp = a.tab_ptr + a.ptr_off[k];
        const u32 *tt = a.tab_term + a.term_off[k];
        const T *u = urow + (u64)(k - 1) * R;
        for (u64 J = threadIdx.x; J < a.Pk[k]; J += GEN_BLOCK) {
          T s = 0;
          for (u32 q = tp[J]; q < tp[J + 1]; ++q) {
            const u32 e = tt[q];
            s += prev[e >> 8] * u[e & 255u];
          }
          dst[J] = s;
        }
        __syncthreads();
        prev = dst;
      }
      {
        const u32 *tp = a.tab_ptr + a.ptr_off[m];
        const u32 *tt = a.tab_term + a.term_off[m];
        const T *u = urow + (u64)(m - 1) * R;
        const T c = coef[o];
        for (u64 J = threadIdx.x; J < a.P; J += GEN_BLOCK) {
          T s = 0;
          for (u32 q = tp[J]; q < tp[J + 1]; ++q) {
            const u32 e = tt[q];
            s += prev[e >> 8] * u[e & 255u];
          }
          acc[J] += c * s;
        }
      }
    }
    double *out = (ch.partial >= 0) ? a.partial + (u64)ch.partial * a.P
                                    : a.Y + (u64)r * a.P;
    for (u64 J = threadIdx.x; J < a.P; J += GEN_BLOCK)
      out[J] = (double)acc[J] * a.fact[J];
  }
}

// grid: (split rows, column blocks of 64), 512 threads: wave w sums the
// w-th eighth of the row's pieces (four interleaved chains), then the eight
// wave sums are added in wave order -- a fixed summation order.
constexpr int kSplitWaves = 8;

__global__ void __launch_bounds__(64 * kSplitWaves)
reduce_split_rows(const SplitRow *__restrict__ split, const double *__restrict__ partial, u64 P,
                  double *__restrict__ Y) {
  __shared__ double red[kSplitWaves][64];
  const SplitRow s = split[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const u64 J = (u64)blockIdx.y * 64 + lane;
  const u64 n = s.slot_end - s.slot_begin;
  const u32 q0 = s.slot_begin + (u32)(n * w / kSplitWaves);
  const u32 q1 = s.slot_begin + (u32)(n * (w + 1) / kSplitWaves);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  if (J < P) {
    u32 q = q0;
    for (; q + 4 <= q1; q += 4) {
      v0 += partial[(u64)q * P + J];
      v1 += partial[(u64)(q + 1) * P + J];
      v2 += partial[(u64)(q + 2) * P + J];
      v3 += partial[(u64)(q + 3) * P + J];
    }
    for (; q < q1; ++q) v0 += partial[(u64)q * P + J];
  }
  red[w][lane] = (v0 + v1) + (v2 + v3);
  __syncthreads();
  if (w == 0 && J < P) {
    double t = red[0][lane];
    for (int i = 1; i < kSplitWaves; ++i) t += red[i][lane];
    Y[(u64)s.row * P + J] = t;
  }
}

template <class T>
__global__ void convert_factor(const double *__restrict__ U, u64 n, T *__restrict__ out) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)U[i];
}

// Batches of up to `bo` owners of one row for the MFMA kernel.
void push_batches(std::vector<Batch> &out, u32 row, u64 ob, u64 oe, u32 bo) {
  auto mk = [&](u64 o, u32 nb, bool last) {
    out.push_back(Batch{(u32)o, (u32)(o >> 32), row, nb | (last ? 0x80000000u : 0u)});
  };
  if (ob == oe) {
    mk(ob, 0, true);
    return;
  }
  for (u64 o = ob; o < oe; o += bo) {
    const u64 e = std::min(oe, o + bo);
    mk(o, (u32)(e - o), e == oe);
  }
}

void build_chunks(Ctx &c, Plan &pl, u64 target, u32 bo) {
  pl.h_chunks.clear();
  pl.h_split.clear();
  pl.h_batches.clear();
  const auto &sp = c.h_segptr;
  const u32 n = c.dim;
  u64 slot = 0;
  constexpr u32 MAX_ROWS = 64;
  u32 r = 0;
  while (r < n) {
    const u64 cnt = sp[r + 1] - sp[r];
    if (cnt > target) {
      SplitRow s{r, (u32)slot, 0};
      for (u64 o = sp[r]; o < sp[r + 1]; o += target) {
        const u64 e = std::min(sp[r + 1], o + target);
        const u64 b0 = pl.h_batches.size();
        push_batches(pl.h_batches, r, o, e, bo);
        pl.h_chunks.push_back(Chunk{o, e, r, r + 1, (i64)slot++, b0, pl.h_batches.size()});
      }
      s.slot_end = (u32)slot;
      pl.h_split.push_back(s);
      ++r;
      continue;
    }
    const u32 start = r;
    u64 acc = 0;
    const u64 b0 = pl.h_batches.size();
    while (r < n && r - start < MAX_ROWS) {
      const u64 k = sp[r + 1] - sp[r];
      if (k > target || (acc + k > target && r > start)) break;
      acc += k;
      push_batches(pl.h_batches, r, sp[r], sp[r + 1], bo);
      ++r;
    }
    pl.h_chunks.push_back(Chunk{sp[start], sp[r], start, r, -1, b0, pl.h_batches.size()});
  }
  pl.n_partial = slot;
}


void upload_plan(Plan &pl) {
  pl.d_chunks.alloc(std::max<size_t>(1, pl.h_chunks.size()) * sizeof(Chunk));
  SP_HIP(hipMemcpy(pl.d_chunks.p, pl.h_chunks.data(), pl.h_chunks.size() * sizeof(Chunk),
                   hipMemcpyHostToDevice));
  pl.d_batches.alloc(std::max<size_t>(1, pl.h_batches.size()) * sizeof(Batch));
  SP_HIP(hipMemcpy(pl.d_batches.p, pl.h_batches.data(), pl.h_batches.size() * sizeof(Batch),
                   hipMemcpyHostToDevice));
  pl.d_split.alloc(std::max<size_t>(1, pl.h_split.size()) * sizeof(SplitRow));
  if (!pl.h_split.empty())
    SP_HIP(hipMemcpy(pl.d_split.p, pl.h_split.data(), pl.h_split.size() * sizeof(SplitRow),
                     hipMemcpyHostToDevice));
}

}  // namespace

void s3ttmc_prepare(Ctx &c, u32 R) {
  if (!c.built) throw Error(SP_ERR_STATE, "tree not built");
  if (c.chunk_R == R && !c.plan[0].h_chunks.empty()) return;
  const int m = (int)c.order - 1;
  if (m > 15) throw Error(SP_ERR_UNSUPPORTED, "order above 16");
  if (R == 0 || R > 255) throw Error(SP_ERR_UNSUPPORTED, "rank must be in [1,255]");
  c.tabs.clear();
  c.tabs.resize(m + 1);
  std::vector<u32> ptr, term;
  c.tab_ptr_off.assign(m + 1, 0);
  c.tab_term_off.assign(m + 1, 0);
  for (int k = 1; k <= m; ++k) {
    c.tabs[k] = make_tables(R, k);
    if (k >= 2) {
      c.tab_ptr_off[k] = ptr.size();
      c.tab_term_off[k] = term.size();
      ptr.insert(ptr.end(), c.tabs[k].term_ptr.begin(), c.tabs[k].term_ptr.end());
      term.insert(term.end(), c.tabs[k].term.begin(), c.tabs[k].term.end());
    }
  }
  if (c.tabs[m].P >= (1ull << 24))
    throw Error(SP_ERR_UNSUPPORTED, "packed width above 2^24");
  ptr.push_back(0);
  term.push_back(0);
  c.d_tab_ptr.alloc(ptr.size() * sizeof(u32));
  c.d_tab_term.alloc(term.size() * sizeof(u32));
  c.d_fact.alloc(c.tabs[m].P * sizeof(double));
  SP_HIP(hipMemcpy(c.d_tab_ptr.p, ptr.data(), ptr.size() * sizeof(u32),
                   hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(c.d_tab_term.p, term.data(), term.size() * sizeof(u32),
                   hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(c.d_fact.p, c.tabs[m].fact_prod.data(),
                   c.tabs[m].P * sizeof(double), hipMemcpyHostToDevice));
  c.P = c.tabs[m].P;
  c.d_W.alloc(c.P * sizeof(double));
  SP_HIP(hipMemcpy(c.d_W.p, c.tabs[m].weight.data(), c.P * sizeof(double),
                   hipMemcpyHostToDevice));

  // specialised MFMA kernel for this (order, rank), with its fold table
  c.fast_bo = 64;
  c.fast_kernel = fast_kernel_for(m, R, &c.fast_H, &c.fast_bo, &c.fast_kernel_f32,
                                  &c.fast_kernel_bf16);
  if (c.fast_kernel) {
    std::vector<u32> fptr, fterm, ftab;
    make_fold_table(R, m, c.fast_H, fptr, fterm, ftab, c.fold_len);
    c.d_fold_ptr.alloc(fptr.size() * sizeof(u32));
    c.d_fold_term.alloc(std::max<size_t>(1, fterm.size()) * sizeof(u32));
    c.d_fold_tab.alloc(ftab.size() * sizeof(u32));
    SP_HIP(hipMemcpy(c.d_fold_ptr.p, fptr.data(), fptr.size() * sizeof(u32),
                     hipMemcpyHostToDevice));
    SP_HIP(hipMemcpy(c.d_fold_term.p, fterm.data(), fterm.size() * sizeof(u32),
                     hipMemcpyHostToDevice));
    SP_HIP(hipMemcpy(c.d_fold_tab.p, ftab.data(), ftab.size() * sizeof(u32),
                     hipMemcpyHostToDevice));
  }

  // work chunks: enough chunks to fill the GPU several times over; the
  // 256-thread MFMA kernel takes larger chunks than the one-wave kernel
  const u64 target =
      c.fast_kernel ? std::max<u64>(256, std::min<u64>(4096, c.n_owners / 2048 + 1))
                    : std::max<u64>(16, std::min<u64>(2048, c.n_owners / 8192 + 1));
  build_chunks(c, c.plan[0], target, 64);
  upload_plan(c.plan[0]);
  u64 np = c.plan[0].n_partial;
  c.plan[1] = Plan{};
  if (c.fast_kernel && c.fast_bo < 64) {
    build_chunks(c, c.plan[1], target, (u32)c.fast_bo);
    upload_plan(c.plan[1]);
    np = std::max(np, c.plan[1].n_partial);
  }
  c.d_partial.reserve(std::max<u64>(1, np) * c.P * sizeof(double));
  c.chunk_R = R;
}

namespace {

// The work plan a launch of the given precision uses.
const Plan &plan_for(const Ctx &c, int precision) {
  return (precision == SP_PREC_FP64 && c.fast_kernel && c.fast_bo < 64) ? c.plan[1] : c.plan[0];
}

// Packed widths per chain level and the LDS layout of the generic kernel.
u64 generic_layout(const Ctx &c, GenArgs &a) {
  const int m = (int)c.order - 1;
  u64 off = 0;
  for (int k = 1; k <= m; ++k) {
    a.ptr_off[k] = c.tab_ptr_off[k];
    a.term_off[k] = c.tab_term_off[k];
    a.Pk[k] = c.tabs[k].P;
    if (k >= 2 && k < m) {
      a.chain_off[k] = off;
      off += c.tabs[k].P;
    }
  }
  a.acc_off = off;
  off += c.tabs[m].P;
  a.u_off = off;
  off += (u64)m * c.chunk_R;
  return off;
}

}  // namespace

u64 s3ttmc_partial_slots(const Ctx &c, int precision) {
  return plan_for(c, precision).n_partial;
}

void s3ttmc_launch(Ctx &c, int precision, const double *U, float *U32, double *Y,
                   double *partial, bool timed) {
  const u32 R = c.chunk_R;
  const int m = (int)c.order - 1;
  const u64 nU = (u64)c.dim * R;
  const Plan &pl = plan_for(c, precision);
  const unsigned nchunks = (unsigned)pl.h_chunks.size();
  const bool prof = c.profile;
  if (!timed) c.profile = false;
  struct Restore {
    Ctx &c;
    bool p;
    ~Restore() { c.profile = p; }
  } restore{c, prof};

  if (precision != SP_PREC_FP64) {
    convert_factor<float><<<grid_for(nU, 256), 256, 0, c.stream>>>(U, nU, U32);
    SP_CHECK_LAUNCH();
  }
  if (precision == SP_PREC_FP64 && c.fast_kernel) {
    ScopedTimer t(c, "s3ttmc_mfma64");
    launch_fast(c.fast_kernel, nchunks, c.stream, pl.d_chunks.as<Chunk>(),
                pl.d_batches.as<Batch>(), c.w_tuple.as<u32>(), c.w_coef.as<double>(), U,
                nullptr, nullptr, c.d_fold_ptr.as<u32>(), c.d_fold_term.as<u32>(),
                c.d_fold_tab.as<u32>(), c.fold_len, c.d_fact.as<double>(), Y, partial);
    SP_CHECK_LAUNCH();
  } else if (precision == SP_PREC_FP32 && c.fast_kernel_f32) {
    ScopedTimer t(c, "s3ttmc_mfma_f32");
    launch_fast(c.fast_kernel_f32, nchunks, c.stream, pl.d_chunks.as<Chunk>(),
                pl.d_batches.as<Batch>(), c.w_tuple.as<u32>(), nullptr, nullptr,
                c.w_coef32.as<float>(), U32, c.d_fold_ptr.as<u32>(), c.d_fold_term.as<u32>(),
                c.d_fold_tab.as<u32>(), c.fold_len, c.d_fact.as<double>(), Y, partial);
    SP_CHECK_LAUNCH();
  } else if (precision == SP_PREC_BF16 && c.fast_kernel_bf16) {
    ScopedTimer t(c, "s3ttmc_mfma_bf16");
    launch_fast(c.fast_kernel_bf16, nchunks, c.stream, pl.d_chunks.as<Chunk>(),
                pl.d_batches.as<Batch>(), c.w_tuple.as<u32>(), nullptr, nullptr,
                c.w_coef32.as<float>(), U32, c.d_fold_ptr.as<u32>(), c.d_fold_term.as<u32>(),
                c.d_fold_tab.as<u32>(), c.fold_len, c.d_fact.as<double>(), Y, partial);
    SP_CHECK_LAUNCH();
  } else {
    GenArgs a{};
    a.tup = c.w_tuple.as<u32>();
    a.segptr = c.w_segptr.as<u64>();
    a.tab_ptr = c.d_tab_ptr.as<u32>();
    a.tab_term = c.d_tab_term.as<u32>();
    a.fact = c.d_fact.as<double>();
    const u64 off = generic_layout(c, a);
    a.m = m;
    a.R = R;
    a.P = c.P;
    a.Y = Y;
    a.partial = partial;
    a.chunks = pl.d_chunks.as<Chunk>();
    if (precision == SP_PREC_FP64) {
      const size_t lds = off * sizeof(double);
      if (lds > 65536) throw Error(SP_ERR_UNSUPPORTED, "packed tensors exceed LDS");
      ScopedTimer t(c, "s3ttmc_generic");
      s3ttmc_generic<double><<<nchunks, GEN_BLOCK, lds, c.stream>>>(a, U,
                                                                    c.w_coef.as<double>());
      SP_CHECK_LAUNCH();
    } else {
      const size_t lds = off * sizeof(float);
      if (lds > 65536) throw Error(SP_ERR_UNSUPPORTED, "packed tensors exceed LDS");
      ScopedTimer t(c, "s3ttmc_generic32");
      s3ttmc_generic<float><<<nchunks, GEN_BLOCK, lds, c.stream>>>(a, U32,
                                                                   c.w_coef32.as<float>());
      SP_CHECK_LAUNCH();
    }
  }
  if (!pl.h_split.empty()) {
    ScopedTimer t(c, "reduce_split_rows");
    reduce_split_rows<<<dim3((unsigned)pl.h_split.size(), (unsigned)((c.P + 63) / 64)),
                        64 * kSplitWaves, 0, c.stream>>>(pl.d_split.as<SplitRow>(), partial,
                                                         c.P, Y);
    SP_CHECK_LAUNCH();
  }
}

void s3ttmc_run(Ctx &c, int precision) {
  const u32 R = c.R;
  s3ttmc_prepare(c, R);
  c.d_Y.reserve((u64)c.dim * c.P * sizeof(double));
  if (precision != SP_PREC_FP64) c.d_U32.reserve((u64)c.dim * R * sizeof(float));
  s3ttmc_launch(c, precision, c.d_U.as<double>(), c.d_U32.as<float>(), c.d_Y.as<double>(),
                c.d_partial.as<double>(), true);
  c.y_R = R;
}

}  // namespace spamd
