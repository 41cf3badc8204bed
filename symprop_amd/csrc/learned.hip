// Learned warm start: node features of a tensor and a small MLP trained and
// evaluated by hand-written HIP kernels (FP32).  No reference counterpart:
// the predicted factor enters through the initialisation hook (hosvd_init,
// src/sparse/spsymtensor.cpp:339-383).
//
// Features (64 per node), computed from the tensor's owner lists:
//   [0]  log(1 + degree)              degree = owners with rem = node
//   [1]  log(1 + distinct partners)   from the level-1 pairs of the tree
//   [2]  log(1 + max pair multiplicity): over the node's partners j, the
//        largest number of (nonzero, position pair) occurrences of {i, j}
//   [3]  log(1 + repeated partners): partners j with multiplicity >= 2
//   [4:] 60-dim sketch: hash-seeded Gaussian node vectors propagated
//        `rounds` times, s'[r] = sum over owners (L, r, v) of v*sum_{a in L} s[a],
//        L2-normalised per row after each round (FP64, stored as FP32).
// MLP: din -> dh -> dh -> dout, ReLU, row-major W[in][out]; loss = mean over
// rows of the squared row error; Adam.  Every reduction has a fixed order.
#include "sp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace spamd {

namespace {

constexpr int SKETCH = 60;
constexpr int NFEAT = 64;

__host__ __device__ inline u64 mix64(u64 x) {
  u64 z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void sketch_init(u32 n, u64 seed, double *__restrict__ s) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (u64)n * SKETCH) return;
  const u64 j = t / SKETCH, k = t % SKETCH;
  const u64 base = seed * 0x100000001B3ull + j * (2 * SKETCH) + 2 * k;
  const u64 a = mix64(base), b = mix64(base + 1);
  const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
  s[t] = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

// One propagation round over the rem-sorted owner lists, in two passes so
// that a hub row (millions of owners in a Zipf tensor) is spread over many
// workgroups:
//   sketch_piece:  workgroup per piece (at most kSketchPiece owners of one
//                  row); wave w takes owners w, w+4, ...; the 4 wave sums are
//                  added in a fixed order into the piece's partial row;
//   sketch_finish: workgroup per row, sums its pieces' partials in piece
//                  order, then L2-normalises the row.
constexpr u64 kSketchPiece = 2048;

struct SketchPiece {
  u64 o0, o1;
};

__global__ void __launch_bounds__(256)
sketch_piece(const SketchPiece *__restrict__ pieces, const u32 *__restrict__ tup,
             const double *__restrict__ val, int m, const double *__restrict__ s,
             double *__restrict__ partial) {
  __shared__ double red[4][64];
  const SketchPiece pc = pieces[blockIdx.x];
  const int k = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc = 0.0;
  if (k < SKETCH) {
    for (u64 o = pc.o0 + w; o < pc.o1; o += 4) {
      double sum = 0.0;
      for (int a = 0; a < m; ++a) sum += s[(u64)tup[o * m + a] * SKETCH + k];
      acc += val[o] * sum;
    }
  }
  red[w][k] = acc;
  __syncthreads();
  if (w == 0 && k < SKETCH)
    partial[(u64)blockIdx.x * SKETCH + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

__global__ void __launch_bounds__(64)
sketch_finish(const u64 *__restrict__ piece_ptr, const double *__restrict__ partial,
              double *__restrict__ out) {
  const u32 r = blockIdx.x;
  const int k = threadIdx.x;
  double acc = 0.0;
  if (k < SKETCH)
    for (u64 p = piece_ptr[r]; p < piece_ptr[r + 1]; ++p) acc += partial[p * SKETCH + k];
  double sq = acc * acc;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) sq += __shfl_xor(sq, w, 64);
  if (k < SKETCH) out[(u64)r * SKETCH + k] = sq > 0.0 ? acc / sqrt(sq) : 0.0;
}

__global__ void gather_rows(const float *__restrict__ src, const long *__restrict__ idx,
                            long rows, int cols, float *__restrict__ dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  const long i = t / cols;
  dst[t] = src[idx[i] * cols + t % cols];
}

__global__ void partner_counts(const u64 *__restrict__ keys, u64 L, int b,
                               unsigned int *__restrict__ cnt) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const u64 k = keys[i];
  const u32 lo = (u32)(k & ((1ull << b) - 1)), hi = (u32)(k >> b);
  if (lo != hi) {
    atomicAdd(&cnt[hi - 1], 1u);
    atomicAdd(&cnt[lo - 1], 1u);
  }
}

// Every position pair (a < b) of every nonzero as a packed key t[a] << b | t[b]
// (1-based indices, sorted tuples); a pair of equal indices gets `dead`,
// the key of (2^b - 1, 2^b - 1), which no pair of distinct indices has.
__global__ void pair_keys(const u32 *__restrict__ inds, u64 nnz, int N, int b, u64 dead,
                          u64 *__restrict__ keys) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const u32 *t = inds + i * N;
  u64 o = i * (u64)(N * (N - 1) / 2);
  for (int a = 0; a < N; ++a)
    for (int c = a + 1; c < N; ++c)
      keys[o++] = t[a] == t[c] ? dead : ((u64)t[a] << b) | t[c];
}

// Runs of equal keys in the sorted pair keys: the first element of each run
// measures the run and raises both nodes' maximum multiplicity and, for a
// run of at least 2, both nodes' count of repeated partners (integer
// atomics: the result does not depend on their order).
__global__ void pair_runs(const u64 *__restrict__ keys, u64 L, int b, u64 dead,
                          unsigned int *__restrict__ maxmult, unsigned int *__restrict__ rep) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const u64 k = keys[i];
  if (k == dead || (i > 0 && keys[i - 1] == k)) return;
  u64 j = i + 1;
  while (j < L && keys[j] == k) ++j;
  const unsigned int m = (unsigned int)min<u64>(j - i, 0xffffffffull);
  const u32 lo = (u32)(k & ((1ull << b) - 1)), hi = (u32)(k >> b);
  atomicMax(&maxmult[hi - 1], m);
  atomicMax(&maxmult[lo - 1], m);
  if (m >= 2) {
    atomicAdd(&rep[hi - 1], 1u);
    atomicAdd(&rep[lo - 1], 1u);
  }
}

__global__ void assemble_features(u32 n, const u64 *__restrict__ segptr,
                                  const unsigned int *__restrict__ partners,
                                  const unsigned int *__restrict__ maxmult,
                                  const unsigned int *__restrict__ rep,
                                  const double *__restrict__ s, float *__restrict__ f) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (u64)n * NFEAT) return;
  const u64 i = t / NFEAT;
  const int k = (int)(t % NFEAT);
  double v;
  if (k == 0) v = log1p((double)(segptr[i + 1] - segptr[i]));
  else if (k == 1) v = log1p((double)partners[i]);
  else if (k == 2) v = log1p((double)maxmult[i]);
  else if (k == 3) v = log1p((double)rep[i]);
  else v = s[i * SKETCH + (k - 4)];
  f[t] = (float)v;
}

// ------------------------------------------------------------------ MLP ---
constexpr int GT = 64, GK = 16;

// C[i][j] = sum_k A(i,k) B(k,j) (+ beta*C), A(i,k) = A[i*ar + k*ac], B(k,j) =
// B[k*br + j*bc], C row-major with leading dimension ldc.  64x64 tiles, fixed
// K order per output (deterministic).
__global__ void __launch_bounds__(256)
gemm_f32(int M, int N, int K, const float *__restrict__ A, long ar, long ac,
         const float *__restrict__ B, long br, long bc, float beta, float *__restrict__ C,
         long ldc) {
  __shared__ float sa[GK][GT + 1];
  __shared__ float sb[GK][GT + 1];
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int t = threadIdx.x; t < GK * GT; t += 256) {
      const int kk = t / GT, c = t % GT;
      const int i = i0 + c, j = j0 + c, k = k0 + kk;
      sa[kk][c] = (i < M && k < K) ? A[(long)i * ar + (long)k * ac] : 0.f;
      sb[kk][c] = (j < N && k < K) ? B[(long)k * br + (long)j * bc] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = sa[kk][ty + 16 * u];
        b[u] = sb[kk][tx + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fmaf(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
      if (i < M && j < N) {
        float *c = C + (long)i * ldc + j;
        *c = beta != 0.f ? acc[u][v] + beta * *c : acc[u][v];
      }
    }
}

__global__ void bias_act(float *__restrict__ Z, const float *__restrict__ b, long rows,
                         int cols, int relu) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  float v = Z[t] + b[t % cols];
  Z[t] = relu ? fmaxf(v, 0.f) : v;
}

// D = 2 (O - T) / N; partial loss sums per block (fixed order later)
__global__ void __launch_bounds__(256)
loss_diff(const float *__restrict__ O, const float *__restrict__ T, long n_el, float scale,
          float *__restrict__ D, double *__restrict__ part) {
  __shared__ double red[256];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (t < n_el) {
    const float d = O[t] - T[t];
    D[t] = 2.f * d * scale;
    sq = (double)d * d;
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void mask_relu(float *__restrict__ dH, const float *__restrict__ H, long n_el) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_el && !(H[t] > 0.f)) dH[t] = 0.f;
}

// column sums: part[blk][j] over a block of rows, then a fixed-order sum
__global__ void colsum_part(const float *__restrict__ M, long rows, int cols, long rows_per,
                            float *__restrict__ part) {
  const int j = threadIdx.x + blockIdx.y * blockDim.x;
  if (j >= cols) return;
  const long r0 = (long)blockIdx.x * rows_per, r1 = min(rows, r0 + rows_per);
  float s = 0.f;
  for (long r = r0; r < r1; ++r) s += M[r * cols + j];
  part[(long)blockIdx.x * cols + j] = s;
}

__global__ void colsum_final(const float *__restrict__ part, int nb, int cols,
                             float *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(long)b * cols + j];
  out[j] = s;
}

__global__ void adam_update(float *__restrict__ p, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, long n, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

unsigned g1(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct Mlp {
  int din, dh, dout;
  hipStream_t stream = nullptr;
  DevBuf params, grad, m, v;
  long off[7];
  // activations for a batch of up to `cap` rows
  long cap = 0;
  DevBuf X, T, H1, H2, O, D, dH2, dH1, part, lpart;
  long step = 0;

  long nparams() const { return off[6]; }
  float *W(int i) const { return params.as<float>() + off[i]; }
  float *G(int i) const { return grad.as<float>() + off[i]; }

  void ensure(long rows) {
    if (rows <= cap) return;
    cap = rows;
    X.alloc(rows * din * sizeof(float));
    T.alloc(rows * dout * sizeof(float));
    H1.alloc(rows * dh * sizeof(float));
    H2.alloc(rows * dh * sizeof(float));
    O.alloc(rows * dout * sizeof(float));
    D.alloc(rows * dout * sizeof(float));
    dH2.alloc(rows * dh * sizeof(float));
    dH1.alloc(rows * dh * sizeof(float));
    lpart.alloc((g1(rows * dout) + 1) * sizeof(double));
  }

  void gemm(int M, int N, int K, const float *A, long ar, long ac, const float *B, long br,
            long bc, float beta, float *C, long ldc) {
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    gemm_f32<<<grid, 256, 0, stream>>>(M, N, K, A, ar, ac, B, br, bc, beta, C, ldc);
    SP_CHECK_LAUNCH();
  }

  void colsum(const float *Mx, long rows, int cols, float *out) {
    const long rows_per = 256;
    const int nb = (int)std::max<long>(1, (rows + rows_per - 1) / rows_per);
    part.reserve((size_t)nb * cols * sizeof(float));
    colsum_part<<<dim3(nb, (cols + 255) / 256), 256, 0, stream>>>(Mx, rows, cols, rows_per,
                                                                   part.as<float>());
    SP_CHECK_LAUNCH();
    colsum_final<<<(cols + 255) / 256, 256, 0, stream>>>(part.as<float>(), nb, cols, out);
    SP_CHECK_LAUNCH();
  }

  // forward on X (rows x din, device) -> O
  void forward(long rows) {
    gemm((int)rows, dh, din, X.as<float>(), din, 1, W(0), dh, 1, 0.f, H1.as<float>(), dh);
    bias_act<<<g1(rows * dh), 256, 0, stream>>>(H1.as<float>(), W(1), rows, dh, 1);
    gemm((int)rows, dh, dh, H1.as<float>(), dh, 1, W(2), dh, 1, 0.f, H2.as<float>(), dh);
    bias_act<<<g1(rows * dh), 256, 0, stream>>>(H2.as<float>(), W(3), rows, dh, 1);
    gemm((int)rows, dout, dh, H2.as<float>(), dh, 1, W(4), dout, 1, 0.f, O.as<float>(), dout);
    bias_act<<<g1(rows * dout), 256, 0, stream>>>(O.as<float>(), W(5), rows, dout, 0);
    SP_CHECK_LAUNCH();
  }

  // loss and gradient of the batch in X/T; returns the loss
  double backward(long rows) {
    const long nel = rows * dout;
    const unsigned nb = g1(nel);
    loss_diff<<<nb, 256, 0, stream>>>(O.as<float>(), T.as<float>(), nel, 1.f / (float)rows,
                                      D.as<float>(), lpart.as<double>());
    SP_CHECK_LAUNCH();
    // W3, b3
    gemm(dh, dout, (int)rows, H2.as<float>(), 1, dh, D.as<float>(), dout, 1, 0.f, G(4), dout);
    colsum(D.as<float>(), rows, dout, G(5));
    // dH2 = D W3^T, masked
    gemm((int)rows, dh, dout, D.as<float>(), dout, 1, W(4), 1, dout, 0.f, dH2.as<float>(), dh);
    mask_relu<<<g1(rows * dh), 256, 0, stream>>>(dH2.as<float>(), H2.as<float>(), rows * dh);
    gemm(dh, dh, (int)rows, H1.as<float>(), 1, dh, dH2.as<float>(), dh, 1, 0.f, G(2), dh);
    colsum(dH2.as<float>(), rows, dh, G(3));
    gemm((int)rows, dh, dh, dH2.as<float>(), dh, 1, W(2), 1, dh, 0.f, dH1.as<float>(), dh);
    mask_relu<<<g1(rows * dh), 256, 0, stream>>>(dH1.as<float>(), H1.as<float>(), rows * dh);
    gemm(din, dh, (int)rows, X.as<float>(), 1, din, dH1.as<float>(), dh, 1, 0.f, G(0), dh);
    colsum(dH1.as<float>(), rows, dh, G(1));
    std::vector<double> lp(nb);
    SP_HIP(hipMemcpyAsync(lp.data(), lpart.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                          stream));
    SP_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double x : lp) s += x;
    return s / (double)rows;
  }
};

Mlp *mlp_create(int din, int dh, int dout, u64 seed) {
  auto *m = new Mlp();
  m->din = din;
  m->dh = dh;
  m->dout = dout;
  const long sizes[6] = {(long)din * dh, dh, (long)dh * dh, dh, (long)dh * dout, dout};
  m->off[0] = 0;
  for (int i = 0; i < 6; ++i) m->off[i + 1] = m->off[i] + sizes[i];
  SP_HIP(hipStreamCreate(&m->stream));
  const long np = m->nparams();
  std::vector<float> p(np, 0.f);
  // He-normal weights (splitmix64 + Box-Muller), zero biases
  const int fan_in[3] = {din, dh, dh};
  for (int l = 0; l < 3; ++l) {
    const double sd = std::sqrt(2.0 / fan_in[l]);
    for (long i = 0; i < sizes[2 * l]; ++i) {
      const u64 base = seed * 0x2545F4914F6CDD1Dull + (u64)(m->off[2 * l] + i) * 2;
      const u64 a = mix64(base), b = mix64(base + 1);
      const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
      const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
      p[m->off[2 * l] + i] =
          (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
    }
  }
  m->params.alloc(np * sizeof(float));
  m->grad.alloc(np * sizeof(float));
  m->m.alloc(np * sizeof(float));
  m->v.alloc(np * sizeof(float));
  SP_HIP(hipMemcpy(m->params.p, p.data(), np * sizeof(float), hipMemcpyHostToDevice));
  SP_HIP(hipMemset(m->m.p, 0, np * sizeof(float)));
  SP_HIP(hipMemset(m->v.p, 0, np * sizeof(float)));
  return m;
}

void mlp_destroy(Mlp *m) {
  if (!m) return;
  hipStream_t s = m->stream;
  if (s) (void)hipStreamSynchronize(s);
  delete m;
  if (s) (void)hipStreamDestroy(s);
}

long mlp_num_params(Mlp *m) { return m->nparams(); }

void mlp_get_params(Mlp *m, float *out) {
  SP_HIP(hipMemcpy(out, m->params.p, m->nparams() * sizeof(float), hipMemcpyDeviceToHost));
}

void mlp_set_params(Mlp *m, const float *in) {
  SP_HIP(hipMemcpy(m->params.p, in, m->nparams() * sizeof(float), hipMemcpyHostToDevice));
  SP_HIP(hipMemset(m->m.p, 0, m->nparams() * sizeof(float)));
  SP_HIP(hipMemset(m->v.p, 0, m->nparams() * sizeof(float)));
  m->step = 0;
}

double mlp_loss_grad(Mlp *m, const float *X, const float *T, long rows, float *grad) {
  m->ensure(rows);
  SP_HIP(hipMemcpyAsync(m->X.p, X, rows * m->din * sizeof(float), hipMemcpyHostToDevice,
                        m->stream));
  SP_HIP(hipMemcpyAsync(m->T.p, T, rows * m->dout * sizeof(float), hipMemcpyHostToDevice,
                        m->stream));
  m->forward(rows);
  const double L = m->backward(rows);
  if (grad)
    SP_HIP(hipMemcpy(grad, m->grad.p, m->nparams() * sizeof(float), hipMemcpyDeviceToHost));
  return L;
}

// Minibatch Adam over the rows (batches are contiguous slices of a seeded
// row permutation drawn on the host each epoch).  Returns the last loss.
double mlp_train(Mlp *m, const float *X, const float *T, long rows, int steps, long batch,
                 float lr, u64 seed, float *losses) {
  batch = std::max<long>(1, std::min(batch, rows));
  DevBuf dX(rows * m->din * sizeof(float)), dT(rows * m->dout * sizeof(float));
  SP_HIP(hipMemcpy(dX.p, X, rows * m->din * sizeof(float), hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(dT.p, T, rows * m->dout * sizeof(float), hipMemcpyHostToDevice));
  m->ensure(batch);
  std::vector<long> perm(rows);
  for (long i = 0; i < rows; ++i) perm[i] = i;
  long pos = rows;
  u64 rs = seed;
  double L = 0.0;
  const long np = m->nparams();
  DevBuf didx(batch * sizeof(long));
  for (int s = 0; s < steps; ++s) {
    if (pos + batch > rows) {  // new epoch: Fisher-Yates with splitmix64
      for (long i = rows - 1; i > 0; --i) {
        rs = mix64(rs);
        std::swap(perm[i], perm[(long)(rs % (u64)(i + 1))]);
      }
      pos = 0;
    }
    // gather the batch rows on the device
    SP_HIP(hipMemcpyAsync(didx.p, perm.data() + pos, batch * sizeof(long),
                          hipMemcpyHostToDevice, m->stream));
    pos += batch;
    gather_rows<<<g1(batch * m->din), 256, 0, m->stream>>>(dX.as<float>(), didx.as<long>(),
                                                           batch, m->din, m->X.as<float>());
    gather_rows<<<g1(batch * m->dout), 256, 0, m->stream>>>(dT.as<float>(), didx.as<long>(),
                                                            batch, m->dout, m->T.as<float>());
    SP_CHECK_LAUNCH();
    m->forward(batch);
    L = m->backward(batch);
    if (losses) losses[s] = (float)L;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    adam_update<<<g1(np), 256, 0, m->stream>>>(m->params.as<float>(), m->grad.as<float>(),
                                              m->m.as<float>(), m->v.as<float>(), np, lr, b1,
                                              b2, c1, c2, 1e-8f);
    SP_CHECK_LAUNCH();
  }
  SP_HIP(hipStreamSynchronize(m->stream));
  return L;
}

void mlp_predict(Mlp *m, const float *X, long rows, float *out) {
  m->ensure(rows);
  SP_HIP(hipMemcpyAsync(m->X.p, X, rows * m->din * sizeof(float), hipMemcpyHostToDevice,
                        m->stream));
  m->forward(rows);
  SP_HIP(hipMemcpyAsync(out, m->O.p, rows * m->dout * sizeof(float), hipMemcpyDeviceToHost,
                        m->stream));
  SP_HIP(hipStreamSynchronize(m->stream));
}

// Node features of the tensor in context c into f (n x 64, host).
void node_features(Ctx &c, int rounds, u64 seed, float *f) {
  if (!c.built) throw Error(SP_ERR_STATE, "tree not built");
  const u32 n = c.dim;
  const int m = (int)c.order - 1;
  const u64 no = c.n_owners;
  (void)no;
  DevBuf s0((u64)n * SKETCH * sizeof(double)), s1((u64)n * SKETCH * sizeof(double));
  sketch_init<<<grid_for((u64)n * SKETCH, 256), 256, 0, c.stream>>>(n, seed, s0.as<double>());
  SP_CHECK_LAUNCH();
  // pieces of at most kSketchPiece owners, row by row
  std::vector<SketchPiece> pieces;
  std::vector<u64> piece_ptr(n + 1);
  const auto &sp = c.h_segptr;
  for (u32 r = 0; r < n; ++r) {
    piece_ptr[r] = pieces.size();
    for (u64 o = sp[r]; o < sp[r + 1]; o += kSketchPiece)
      pieces.push_back(SketchPiece{o, std::min(sp[r + 1], o + kSketchPiece)});
  }
  piece_ptr[n] = pieces.size();
  const u64 np = pieces.size();
  if (np >= (1ull << 31)) throw Error(SP_ERR_UNSUPPORTED, "too many sketch pieces");
  DevBuf dpieces(std::max<u64>(1, np) * sizeof(SketchPiece));
  DevBuf dpp((u64)(n + 1) * sizeof(u64));
  DevBuf partial(std::max<u64>(1, np) * SKETCH * sizeof(double));
  if (np)
    SP_HIP(hipMemcpyAsync(dpieces.p, pieces.data(), np * sizeof(SketchPiece),
                          hipMemcpyHostToDevice, c.stream));
  SP_HIP(hipMemcpyAsync(dpp.p, piece_ptr.data(), (u64)(n + 1) * sizeof(u64),
                        hipMemcpyHostToDevice, c.stream));
  for (int r = 0; r < rounds; ++r) {
    if (np) {
      sketch_piece<<<(unsigned)np, 256, 0, c.stream>>>(
          dpieces.as<SketchPiece>(), c.w_tuple.as<u32>(), c.w_val.as<double>(), m,
          s0.as<double>(), partial.as<double>());
      SP_CHECK_LAUNCH();
    }
    sketch_finish<<<n, 64, 0, c.stream>>>(dpp.as<u64>(), partial.as<double>(),
                                          s1.as<double>());
    SP_CHECK_LAUNCH();
    std::swap(s0, s1);
  }
  // the host vectors above are read by the async copies: finish them here
  SP_HIP(hipStreamSynchronize(c.stream));
  DevBuf partners((u64)n * sizeof(unsigned int));
  SP_HIP(hipMemsetAsync(partners.p, 0, (u64)n * sizeof(unsigned int), c.stream));
  if (c.depth >= 2) {
    if (c.depth > 2) cdtree_ensure_levels(c);
    // level 1 holds pairs: 2 * bits(dim) <= 64, so its keys are u64
    const DevLevel &l1 = c.levels[1];
    partner_counts<<<grid_for(l1.size, 256), 256, 0, c.stream>>>(
        l1.keys.as<u64>(), l1.size, c.bits, partners.as<unsigned int>());
    SP_CHECK_LAUNCH();
  }
  // pair multiplicities: sort every position pair of every nonzero
  DevBuf maxmult((u64)n * sizeof(unsigned int)), rep((u64)n * sizeof(unsigned int));
  SP_HIP(hipMemsetAsync(maxmult.p, 0, (u64)n * sizeof(unsigned int), c.stream));
  SP_HIP(hipMemsetAsync(rep.p, 0, (u64)n * sizeof(unsigned int), c.stream));
  {
    const int N = (int)c.order, b = c.bits;
    if (2 * b > 64) throw Error(SP_ERR_UNSUPPORTED, "pair keys need 2*bits(dim) <= 64");
    const u64 np2 = c.nnz * (u64)(N * (N - 1) / 2);
    const u64 mask = (b >= 64) ? ~0ull : ((1ull << b) - 1);
    const u64 dead = (mask << b) | mask;
    DevBuf ka(np2 * sizeof(u64)), kb(np2 * sizeof(u64));
    pair_keys<<<grid_for(c.nnz, 256), 256, 0, c.stream>>>(c.d_inds.as<u32>(), c.nnz, N, b,
                                                          dead, ka.as<u64>());
    SP_CHECK_LAUNCH();
    u64 *sorted = nullptr;
    u32 *unused = nullptr;
    radix_sort<u64>(ka.as<u64>(), kb.as<u64>(), nullptr, nullptr, np2, 2 * b, c.stream,
                    &sorted, &unused);
    pair_runs<<<grid_for(np2, 256), 256, 0, c.stream>>>(sorted, np2, b, dead,
                                                       maxmult.as<unsigned int>(),
                                                       rep.as<unsigned int>());
    SP_CHECK_LAUNCH();
    SP_HIP(hipStreamSynchronize(c.stream));
  }
  DevBuf df((u64)n * NFEAT * sizeof(float));
  assemble_features<<<grid_for((u64)n * NFEAT, 256), 256, 0, c.stream>>>(
      n, c.w_segptr.as<u64>(), partners.as<unsigned int>(), maxmult.as<unsigned int>(),
      rep.as<unsigned int>(), s0.as<double>(), df.as<float>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipMemcpyAsync(f, df.p, (u64)n * NFEAT * sizeof(float), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace spamd
