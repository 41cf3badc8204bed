// C-ABI entry points (declared in include/symprop_amd.h).
#include "sp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

using namespace spamd;

struct sp_ctx {
  Ctx c;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guard(sp_ctx *ctx, F &&f) {
  try {
    f();
    return SP_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    if (ctx) ctx->c.last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return SP_ERR_OOM;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return SP_ERR_DEVICE;
  }
}

void need(bool cond, int code, const char *msg) {
  if (!cond) throw Error(code, msg);
}

void need_factor(Ctx &c) {
  need(c.R > 0 && c.d_U.p, SP_ERR_STATE, "no factor uploaded");
}

void need_y(Ctx &c) {
  need(c.d_Y.p && c.P > 0 && c.y_R != 0, SP_ERR_STATE, "no S3TTMc result on the device");
  need(c.y_R == c.R, SP_ERR_STATE,
       "the device S3TTMc result was computed for another rank (run S3TTMc again)");
}

void need_core(Ctx &c) {
  need(c.d_core.p && c.core_R != 0, SP_ERR_STATE, "no core (call sp_gemmtn first)");
  need(c.core_R == c.R, SP_ERR_STATE,
       "the device core was computed for another rank (call sp_gemmtn again)");
}

}  // namespace

namespace spamd {
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace spamd

extern "C" {

const char *sp_last_error(void) { return g_last_error.c_str(); }

const char *sp_version(void) { return "symprop_amd 0.1 (gfx950)"; }

int sp_ctx_create(uint32_t order, uint32_t dim, uint64_t nnz, const uint32_t *inds,
                  const double *vals, sp_ctx **out) {
  return guard(nullptr, [&] {
    need(out != nullptr, SP_ERR_ARG, "out is null");
    *out = nullptr;
    need(inds && vals, SP_ERR_ARG, "null tensor buffers");
    need(order >= 3 && order <= 16, SP_ERR_UNSUPPORTED, "order must be in [3,16]");
    need(dim >= 1, SP_ERR_ARG, "dim must be >= 1");
    need(nnz >= 1, SP_ERR_ARG, "nnz must be >= 1");
    int ndev = 0;
    SP_HIP(hipGetDeviceCount(&ndev));
    need(ndev > 0, SP_ERR_DEVICE, "no HIP device");
    auto *h = new sp_ctx();
    Ctx &c = h->c;
    try {
      c.order = order;
      c.dim = dim;
      c.nnz = nnz;
      if (const char *e = std::getenv("SP_TRD_SPIN"))
        c.trd_spin_limit = (unsigned)std::max(1L, std::strtol(e, nullptr, 10));
      if (const char *e = std::getenv("SP_EIG_WARM")) c.warm_enabled = std::strcmp(e, "0") != 0;
      if (const char *e = std::getenv("SP_SYTRD"))
        c.trd_mode = std::strcmp(e, "wg") == 0 ? 1 : std::strcmp(e, "reg") == 0 ? 2 : 0;
      if (const char *e = std::getenv("SP_EIG"))
        c.eig_method = std::strcmp(e, "syevd") == 0               ? 0
                       : std::strcmp(e, "syevj") == 0              ? 1
                       : std::strcmp(e, "tridiag-rocsolver") == 0  ? 3
                                                                   : 2;
      if (const char *e = std::getenv("SP_HOEVD"))
        c.hoevd_method = std::strcmp(e, "dense") == 0 ? 1 : std::strcmp(e, "krylov") == 0 ? 2 : 0;
      SP_HIP(hipGetDevice(&c.device));
      noalloc_setup(c.device);
      SP_HIP(hipStreamCreate(&c.stream));
      c.d_inds.alloc(nnz * order * sizeof(uint32_t));
      c.d_vals.alloc(nnz * sizeof(double));
      SP_HIP(hipMemcpyAsync(c.d_inds.p, inds, nnz * order * sizeof(uint32_t),
                            hipMemcpyHostToDevice, c.stream));
      SP_HIP(hipMemcpyAsync(c.d_vals.p, vals, nnz * sizeof(double), hipMemcpyHostToDevice,
                            c.stream));
      // index range and order checks and the norm, on the device
      input_check_norm(c);
    } catch (...) {
      if (c.stream) (void)hipStreamSynchronize(c.stream);
      hipStream_t s = c.stream;
      delete h;
      if (s) (void)hipStreamDestroy(s);
      throw;
    }
    *out = h;
  });
}

int sp_ctx_destroy(sp_ctx *ctx) {
  if (!ctx) return SP_OK;
  return guard(nullptr, [&] {
    Ctx &c = ctx->c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (auto &t : c.timers) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    hipStream_t s = c.stream;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
  });
}

int sp_ctx_info(sp_ctx *ctx, uint32_t *order, uint32_t *dim, uint64_t *nnz) {
  return guard(ctx, [&] {
    need(ctx && order && dim && nnz, SP_ERR_ARG, "null argument");
    *order = ctx->c.order;
    *dim = ctx->c.dim;
    *nnz = ctx->c.nnz;
  });
}

int sp_norm(sp_ctx *ctx, double *out) {
  return guard(ctx, [&] {
    need(ctx && out, SP_ERR_ARG, "null argument");
    *out = ctx->c.norm;
  });
}

int sp_cdtree_build(sp_ctx *ctx) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    ScopedTimer t(ctx->c, "cdtree_build");
    cdtree_build(ctx->c);
  });
}

int sp_cdtree_depth(sp_ctx *ctx, uint32_t *depth) {
  return guard(ctx, [&] {
    need(ctx && depth, SP_ERR_ARG, "null argument");
    need(ctx->c.built, SP_ERR_STATE, "tree not built");
    *depth = (uint32_t)ctx->c.depth;
  });
}

int sp_cdtree_level_size(sp_ctx *ctx, uint32_t level, uint64_t *size) {
  return guard(ctx, [&] {
    need(ctx && size, SP_ERR_ARG, "null argument");
    need(ctx->c.built, SP_ERR_STATE, "tree not built");
    need(level < (uint32_t)ctx->c.depth, SP_ERR_ARG, "level out of range");
    if ((int)level < ctx->c.depth - 1) cdtree_ensure_levels(ctx->c);
    *size = ctx->c.levels[level].size;
  });
}

int sp_cdtree_num_owners(sp_ctx *ctx, uint64_t *count) {
  return guard(ctx, [&] {
    need(ctx && count, SP_ERR_ARG, "null argument");
    need(ctx->c.built, SP_ERR_STATE, "tree not built");
    *count = ctx->c.n_owners;
  });
}

int sp_cdtree_export(sp_ctx *ctx, uint32_t level, uint32_t *keys, uint32_t *level_idx,
                     uint64_t *level_ptr, uint64_t *dep_list) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    need(ctx->c.built, SP_ERR_STATE, "tree not built");
    need(level < (uint32_t)ctx->c.depth, SP_ERR_ARG, "level out of range");
    cdtree_export_level(ctx->c, (int)level, keys, level_idx, level_ptr, dep_list);
  });
}

int sp_cdtree_export_owners(sp_ctx *ctx, uint32_t *rem, double *val) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    need(ctx->c.built, SP_ERR_STATE, "tree not built");
    cdtree_export_owners(ctx->c, rem, val);
  });
}

int sp_factor_upload(sp_ctx *ctx, const double *U, uint32_t R) {
  return guard(ctx, [&] {
    need(ctx && U, SP_ERR_ARG, "null argument");
    need(R >= 1 && R <= 255, SP_ERR_UNSUPPORTED, "rank must be in [1,255]");
    Ctx &c = ctx->c;
    if (c.R != R) {
      c.R = R;
      c.P = symsize(R, c.order - 1);
      c.chunk_R = 0;
    }
    c.d_U.reserve((u64)c.dim * R * sizeof(double));
    SP_HIP(hipMemcpyAsync(c.d_U.p, U, (u64)c.dim * R * sizeof(double),
                          hipMemcpyHostToDevice, c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
  });
}

int sp_factor_download(sp_ctx *ctx, double *U) {
  return guard(ctx, [&] {
    need(ctx && U, SP_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    need_factor(c);
    SP_HIP(hipMemcpyAsync(U, c.d_U.p, (u64)c.dim * c.R * sizeof(double),
                          hipMemcpyDeviceToHost, c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
  });
}

int sp_s3ttmc_dev(sp_ctx *ctx, int precision) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    need(c.built, SP_ERR_STATE, "tree not built");
    need_factor(c);
    need(precision >= SP_PREC_FP64 && precision <= SP_PREC_BF16, SP_ERR_ARG,
         "unknown precision");
    s3ttmc_run(c, precision);
  });
}

int sp_s3ttmc_noalloc(sp_ctx *ctx, const double *dU, uint32_t R, double *dY,
                      int precision) {
  return guard(ctx, [&] {
    need(ctx && dU && dY, SP_ERR_ARG, "null argument");
    need(precision >= SP_PREC_FP64 && precision <= SP_PREC_BF16, SP_ERR_ARG,
         "unknown precision");
    s3ttmc_noalloc(ctx->c, dU, R, dY, precision);
  });
}

int sp_s3ttmc(sp_ctx *ctx, const double *U, uint32_t R, double *Y, uint64_t ldy,
              int precision) {
  int rc = sp_factor_upload(ctx, U, R);
  if (rc) return rc;
  rc = sp_s3ttmc_dev(ctx, precision);
  if (rc) return rc;
  return sp_y_download(ctx, Y, ldy);
}

int sp_y_download(sp_ctx *ctx, double *Y, uint64_t ldy) {
  return guard(ctx, [&] {
    need(ctx && Y, SP_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    need_y(c);
    need(ldy >= c.P, SP_ERR_ARG, "ldy smaller than the packed width");
    SP_HIP(hipMemcpy2DAsync(Y, ldy * sizeof(double), c.d_Y.p, c.P * sizeof(double),
                            c.P * sizeof(double), c.dim, hipMemcpyDeviceToHost,
                            c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
  });
}

int sp_packed_width(sp_ctx *ctx, uint64_t *P) {
  return guard(ctx, [&] {
    need(ctx && P, SP_ERR_ARG, "null argument");
    need_factor(ctx->c);
    *P = ctx->c.P;
  });
}

int sp_gemmtn(sp_ctx *ctx, double *core) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    need_factor(c);
    need_y(c);
    dense_gemmtn(c);
    if (core) {
      SP_HIP(hipMemcpyAsync(core, c.d_core.p, (u64)c.R * c.P * sizeof(double),
                            hipMemcpyDeviceToHost, c.stream));
    }
    SP_HIP(hipStreamSynchronize(c.stream));
  });
}

int sp_symgemm(sp_ctx *ctx) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    need_factor(c);
    need_y(c);
    need_core(c);
    dense_symgemm(c);
  });
}

int sp_gram(sp_ctx *ctx, double *G) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    need_y(c);
    dense_gram(c);
    if (G) {
      SP_HIP(hipMemcpyAsync(G, c.d_G.p, c.P * c.P * sizeof(double),
                            hipMemcpyDeviceToHost, c.stream));
      SP_HIP(hipStreamSynchronize(c.stream));
    }
  });
}

int sp_leftsvd_y(sp_ctx *ctx, double *sigma) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    need_factor(c);
    need_y(c);
    dense_leftsvd_y(c, sigma);
  });
}

int sp_qr(sp_ctx *ctx) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    need_factor(ctx->c);
    need(ctx->c.R <= ctx->c.dim, SP_ERR_ARG, "rank exceeds dimension");
    dense_qr(ctx->c);
  });
}

int sp_regularize(sp_ctx *ctx, double delta) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    need_factor(ctx->c);
    dense_regularize(ctx->c, delta);
  });
}

int sp_core_norm(sp_ctx *ctx, double *out) {
  return guard(ctx, [&] {
    need(ctx && out, SP_ERR_ARG, "null argument");
    need_core(ctx->c);
    *out = dense_core_norm(ctx->c);
  });
}

int sp_factor_snapshot(sp_ctx *ctx) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    need_factor(ctx->c);
    dense_factor_snapshot(ctx->c);
  });
}

int sp_subspace_sin(sp_ctx *ctx, double *out) {
  return guard(ctx, [&] {
    need(ctx && out, SP_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    need_factor(c);
    need(c.ws_uprev.p && c.uprev_R == c.R, SP_ERR_STATE, "no factor snapshot of this rank");
    *out = dense_subspace_sin(c);
  });
}

int sp_hoevd_gram(sp_ctx *ctx, double *A) {
  return guard(ctx, [&] {
    need(ctx && A, SP_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    need(c.built, SP_ERR_STATE, "tree not built");
    const u64 n = c.dim;
    DevBuf d(n * n * sizeof(double));
    hoevd_gram(c, d.as<double>());
    SP_HIP(hipMemcpyAsync(A, d.p, n * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
  });
}

int sp_hoevd_init(sp_ctx *ctx, uint32_t R, double *U, double *eigvals) {
  std::vector<double> Uh;
  const int rc = guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    need(c.built, SP_ERR_STATE, "tree not built");
    need(R >= 1 && R <= 255, SP_ERR_UNSUPPORTED, "rank must be in [1,255]");
    need(R <= c.dim, SP_ERR_ARG, "rank exceeds dimension");
    Uh.resize((u64)c.dim * R);
    hoevd_init(c, R, Uh.data(), eigvals);
    if (U) std::copy(Uh.begin(), Uh.end(), U);
  });
  if (rc) return rc;
  return sp_factor_upload(ctx, Uh.data(), R);
}

int sp_syevd_top(sp_ctx *ctx, const double *A, uint64_t dim, uint32_t rank, double *U,
                 double *w) {
  return guard(ctx, [&] {
    need(ctx && A && U, SP_ERR_ARG, "null argument");
    need(dim >= 1 && dim <= 65535, SP_ERR_ARG, "dim must be in [1,65535]");
    need(rank >= 1 && rank <= dim, SP_ERR_ARG, "rank must be in [1,dim]");
    Ctx &c = ctx->c;
    c.ws_sym.reserve(dim * dim * sizeof(double));
    c.ws_pr.reserve(dim * rank * sizeof(double));
    SP_HIP(hipMemcpyAsync(c.ws_sym.p, A, dim * dim * sizeof(double), hipMemcpyHostToDevice,
                          c.stream));
    std::vector<double> theta, V(dim * rank);
    eig_top(c, c.ws_sym.as<double>(), dim, rank, c.ws_pr.as<double>(), theta);
    SP_HIP(hipMemcpyAsync(V.data(), c.ws_pr.p, dim * rank * sizeof(double),
                          hipMemcpyDeviceToHost, c.stream));
    SP_HIP(hipStreamSynchronize(c.stream));
    // ascending order, as matrix::syevd returns the columns dim-rank..dim-1
    for (u64 i = 0; i < dim; ++i)
      for (u32 j = 0; j < rank; ++j) U[i * rank + j] = V[i * rank + (rank - 1 - j)];
    if (w)
      for (u32 j = 0; j < rank; ++j) w[j] = theta[rank - 1 - j];
  });
}

int sp_node_features(sp_ctx *ctx, uint32_t rounds, uint64_t seed, float *out) {
  return guard(ctx, [&] {
    need(ctx && out, SP_ERR_ARG, "null argument");
    need(rounds <= 64, SP_ERR_ARG, "rounds must be <= 64");
    node_features(ctx->c, (int)rounds, seed, out);
  });
}

int sp_mlp_create(uint32_t din, uint32_t dh, uint32_t dout, uint64_t seed, sp_mlp **out) {
  return guard(nullptr, [&] {
    need(out != nullptr, SP_ERR_ARG, "out is null");
    need(din >= 1 && dh >= 1 && dout >= 1 && din <= 4096 && dh <= 4096 && dout <= 4096,
         SP_ERR_ARG, "layer widths must be in [1,4096]");
    *out = reinterpret_cast<sp_mlp *>(mlp_create((int)din, (int)dh, (int)dout, seed));
  });
}

int sp_mlp_destroy(sp_mlp *m) {
  return guard(nullptr, [&] { mlp_destroy(reinterpret_cast<Mlp *>(m)); });
}

int sp_mlp_num_params(sp_mlp *m, uint64_t *n) {
  return guard(nullptr, [&] {
    need(m && n, SP_ERR_ARG, "null argument");
    *n = (uint64_t)mlp_num_params(reinterpret_cast<Mlp *>(m));
  });
}

int sp_mlp_get_params(sp_mlp *m, float *p) {
  return guard(nullptr, [&] {
    need(m && p, SP_ERR_ARG, "null argument");
    mlp_get_params(reinterpret_cast<Mlp *>(m), p);
  });
}

int sp_mlp_set_params(sp_mlp *m, const float *p) {
  return guard(nullptr, [&] {
    need(m && p, SP_ERR_ARG, "null argument");
    mlp_set_params(reinterpret_cast<Mlp *>(m), p);
  });
}

int sp_mlp_loss_grad(sp_mlp *m, const float *X, const float *T, uint64_t rows, double *loss,
                     float *grad) {
  return guard(nullptr, [&] {
    need(m && X && T && loss && rows >= 1, SP_ERR_ARG, "null argument or empty batch");
    *loss = mlp_loss_grad(reinterpret_cast<Mlp *>(m), X, T, (long)rows, grad);
  });
}

int sp_mlp_train(sp_mlp *m, const float *X, const float *T, uint64_t rows, uint32_t steps,
                 uint64_t batch, float lr, uint64_t seed, float *losses, double *final_loss) {
  return guard(nullptr, [&] {
    need(m && X && T && rows >= 1 && batch >= 1, SP_ERR_ARG, "null argument or empty batch");
    const double L = mlp_train(reinterpret_cast<Mlp *>(m), X, T, (long)rows, (int)steps,
                               (long)batch, lr, seed, losses);
    if (final_loss) *final_loss = L;
  });
}

int sp_mlp_predict(sp_mlp *m, const float *X, uint64_t rows, float *out) {
  return guard(nullptr, [&] {
    need(m && X && out && rows >= 1, SP_ERR_ARG, "null argument or empty batch");
    mlp_predict(reinterpret_cast<Mlp *>(m), X, (long)rows, out);
  });
}

int sp_timing_enable(sp_ctx *ctx, int on) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    ctx->c.profile = on != 0;
  });
}

int sp_timing_get(sp_ctx *ctx, const char *name, double *total_ms, uint64_t *launches) {
  return guard(ctx, [&] {
    need(ctx && name && total_ms && launches, SP_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    SP_HIP(hipStreamSynchronize(c.stream));
    double tot = 0.0;
    uint64_t cnt = 0;
    for (auto &t : c.timers) {
      if (t.name != name) continue;
      float ms = 0.f;
      SP_HIP(hipEventElapsedTime(&ms, t.a, t.b));
      tot += ms;
      ++cnt;
    }
    *total_ms = tot;
    *launches = cnt;
  });
}

int sp_timing_reset(sp_ctx *ctx) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    Ctx &c = ctx->c;
    SP_HIP(hipStreamSynchronize(c.stream));
    for (auto &t : c.timers) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    c.timers.clear();
  });
}

int sp_sync(sp_ctx *ctx) {
  return guard(ctx, [&] {
    need(ctx, SP_ERR_ARG, "null context");
    SP_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

}  // extern "C"
