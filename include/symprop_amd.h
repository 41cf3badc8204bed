/* C-ABI of the MI355X-native S3TTMc hot path for symmetric sparse Tucker.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary.  Every
 * function returns SP_OK (0) or an SP_ERR_* status; sp_last_error() gives
 * the message of the last failure on the calling thread.  Host buffers stay
 * owned by the caller; a context owns its device memory and one HIP stream.
 * One context is used from one host thread at a time; independent contexts
 * may run concurrently.
 *
 * Conventions taken from the reference (paths relative to its checkout):
 *   - tensor indices are 1-based, each nonzero's index tuple sorted
 *     ascending (src/utils/loader.cpp:30), row-major nnz x order uint32;
 *   - factor U is n x R row-major FP64, row i-1 belongs to index i
 *     (src/compressed_dtree.cpp:426);
 *   - packed rows hold the P = C(R+N-2, N-1) entries of a symmetric
 *     order-(N-1) tensor in lexicographic order of sorted multi-indices
 *     (include/detail/dense/symtensor.hpp:129-142).
 */
#ifndef SYMPROP_AMD_H
#define SYMPROP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_OK 0
#define SP_ERR_ARG 1         /* invalid argument (null, size, unsorted tuple) */
#define SP_ERR_UNSUPPORTED 2 /* order / rank / dimension outside the limits   */
#define SP_ERR_DEVICE 3      /* HIP runtime or library failure                */
#define SP_ERR_OOM 4         /* device allocation failed                      */
#define SP_ERR_STATE 5       /* call out of order (no tree, no factor, ...)   */
#define SP_ERR_NUMERIC 6     /* NaN / non-finite result detected              */
#define SP_ERR_IO 7          /* file cannot be opened or parsed               */

#define SP_PREC_FP64 0 /* FP64 everywhere                                    */
#define SP_PREC_FP32 1 /* FP32 values, U rows and accumulation, FP64 Y out   */
#define SP_PREC_BF16 2 /* BF16 operands, FP32 accumulation, FP64 Y out       */

typedef struct sp_ctx sp_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char *sp_last_error(void);

/* Library version string. */
const char *sp_version(void);

/* Uploads the tensor and computes its norm.  Replaces the state set up by
 * SpSymtensor::SpSymtensor (src/sparse/spsymtensor.cpp:16-28, declared at
 * include/detail/sparse/spsymtensor.hpp:28-29); the tree is built by
 * sp_cdtree_build.  Limits: 3 <= order <= 16, nnz >= 1, 1 <= index <= dim. */
int sp_ctx_create(uint32_t order, uint32_t dim, uint64_t nnz,
                  const uint32_t *inds, const double *vals, sp_ctx **out);
int sp_ctx_destroy(sp_ctx *ctx);

/* Shape of the tensor held by a context. */
int sp_ctx_info(sp_ctx *ctx, uint32_t *order, uint32_t *dim, uint64_t *nnz);

/* ||X|| = sqrt(sum v^2 N!/prod f!)  (SpSymtensor::calc_norm,
 * src/sparse/spsymtensor.cpp:114-138). */
int sp_norm(sp_ctx *ctx, double *out);

/* Builds the compressed dimension tree on the GPU: replaces
 * DTree::banerjee_insert (src/dtree.cpp:56-80) + CompressedDTree's
 * constructor and init_dep_list (src/compressed_dtree.cpp:13-131). */
int sp_cdtree_build(sp_ctx *ctx);

/* Tree shape: depth = order-1 levels; level l holds the sorted distinct
 * sub-multisets of size l+1. */
int sp_cdtree_depth(sp_ctx *ctx, uint32_t *depth);
int sp_cdtree_level_size(sp_ctx *ctx, uint32_t level, uint64_t *size);
int sp_cdtree_num_owners(sp_ctx *ctx, uint64_t *count);

/* Copies one level of the tree to host buffers (any may be null):
 *   keys      [size x (level+1)]  full sorted tuples of the nodes
 *   level_idx [size]              last index of each node (level_idx)
 *   level_ptr [size+1]            CSR to level+1, or into the owners at the
 *                                 leaf level (level_ptr)
 *   dep_list  [size x (level+1)]  exclude-one parent indices (dep_list),
 *                                 level >= 1 only
 * matching the members of CompressedDTree
 * (include/detail/compressed_dtree.hpp:52-59) bit for bit. */
int sp_cdtree_export(sp_ctx *ctx, uint32_t level, uint32_t *keys,
                     uint32_t *level_idx, uint64_t *level_ptr,
                     uint64_t *dep_list);
/* Owner list (CompressedDTree::nonzero): (rem index, value) in leaf order. */
int sp_cdtree_export_owners(sp_ctx *ctx, uint32_t *rem, double *val);

/* Uploads / downloads the factor U (n x R).  R may change between calls. */
int sp_factor_upload(sp_ctx *ctx, const double *U, uint32_t R);
int sp_factor_download(sp_ctx *ctx, double *U);

/* Packed S3TTMc of the device factor into the device Y (n x P), overwriting
 * Y.  Replaces SpSymtensor::S3TTMc(res, U, R, partial_sym_res=true)
 * (include/detail/sparse/spsymtensor.hpp:38-39, src/sparse/spsymtensor.cpp:
 * 30-98) and CompressedDTree::kronecker_product_opt5
 * (include/detail/compressed_dtree.hpp:21-22).  precision: SP_PREC_*. */
int sp_s3ttmc_dev(sp_ctx *ctx, int precision);

/* Allocation-free form of sp_s3ttmc_dev.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * the same kernels on the same work plan as sp_s3ttmc_dev (bit-identical
 * result) for the caller's device factor dU (dim x R FP64, row-major) into
 * the caller's device dY (dim x P FP64, row-major, P = C(R+N-2, N-1)),
 * queued on the context's stream (sp_sync waits for it); the context's own
 * factor, Y and core are left alone.  The context must already be prepared
 * for rank R by an sp_s3ttmc_dev call at that rank (SP_ERR_STATE otherwise).
 * Scratch (split-row partial sums, the FP32 copy of U for the reduced
 * precisions) lives in a static 64 MiB device array per GPU that every call
 * on that GPU shares: calls that overlap on one GPU race, and a need beyond
 * 64 MiB is rejected with SP_ERR_UNSUPPORTED.  Calls no hipMalloc/hipFree. */
int sp_s3ttmc_noalloc(sp_ctx *ctx, const double *dU, uint32_t R, double *dY,
                      int precision);

/* Host-buffer convenience form of the same call: uploads U, runs, and copies
 * the packed Y into Y (n rows, leading dimension ldy >= P). */
int sp_s3ttmc(sp_ctx *ctx, const double *U, uint32_t R, double *Y,
              uint64_t ldy, int precision);

/* Copies the device Y (n x P) to the host with leading dimension ldy. */
int sp_y_download(sp_ctx *ctx, double *Y, uint64_t ldy);
/* Packed width P for the current R (C(R+N-2, N-1)). */
int sp_packed_width(sp_ctx *ctx, uint64_t *P);

/* core = U^T Y (R x P) on the device; copied to core (if non-null).
 * Replaces matrix::gemmtn (src/dense/matrixops.cpp:45-56) as called at
 * src/sparse/spsymtensor.cpp:270 and 321. */
int sp_gemmtn(sp_ctx *ctx, double *core);

/* U = Y diag(W) core^T (n x R) on the device, W = symtens_coeff of order
 * N-1.  Replaces matrix::symgemm (src/dense/matrixops.cpp:102-124) as
 * called at src/sparse/spsymtensor.cpp:271. */
int sp_symgemm(sp_ctx *ctx);

/* G = Y^T diag(W) Y (P x P) on the device; copied to G (if non-null).  The
 * n x n Gram Y diag(W) Y^T of matrix::symsyrk (src/dense/matrixops.cpp:
 * 126-146) shares its nonzero spectrum; this is its P x P form. */
int sp_gram(sp_ctx *ctx, double *G);

/* U = top-R left singular vectors of the unpacked Y (n x R^(N-1)), computed
 * from the packed Y through its weighted Gram.  Replaces the
 * matrix::leftsvd(Y) of src/sparse/spsymtensor.cpp:225 and the
 * symsyrk + syevd pair of src/sparse/spsymtensor.cpp:318-319.
 * sigma (if non-null) receives the R singular values, descending. */
int sp_leftsvd_y(sp_ctx *ctx, double *sigma);

/* Thin QR of the device U in place (matrix::qr, src/dense/matrixops.cpp:
 * 58-75, called at src/sparse/spsymtensor.cpp:272). */
int sp_qr(sp_ctx *ctx);

/* Row clipping to norm <= delta, then orthonormalisation of U (the left
 * singular vectors of the clipped U): regularize,
 * src/sparse/spsymtensor.cpp:168-199. */
int sp_regularize(sp_ctx *ctx, double delta);

/* sqrt(sum_r sum_p core[r,p]^2 W[p]) of the device core
 * (symtens_norm, src/dense/symtensor.cpp:97-106). */
int sp_core_norm(sp_ctx *ctx, double *out);

/* Options of one decomposition run: the fields of Config
 * (include/detail/tucker.hpp:14-23) plus this library's extensions.
 * sp_tucker_opts_default fills Config's defaults (rank 2, max_iter 10,
 * rel_tol 1e-6, delta 0.5, seed 0, init "rnd", FP64, automatic switch
 * tolerance, no subspace test, silent). */
typedef struct sp_tucker_opts {
  int alg;             /* 0 TuckerHOOI, 1 TuckerHOQRI, 2 TuckerAltHOOI       */
  uint32_t rank;
  uint64_t max_iter;
  double rel_tol;      /* stop when |error - previous error| < rel_tol     */
  double delta;        /* row clip of regularize                           */
  uint64_t seed;       /* random initial factor (std::mt19937 + N(0,1))    */
  const char *init;    /* "rnd" or "hoevd" (NULL: "rnd"); U0 overrides     */
  int precision;       /* SP_PREC_* of the S3TTMc iterations               */
  double switch_tol;   /* reduced forms: |error change| that ends the
                          reduced phase (<= 0: 1e-3 BF16, 1e-5 FP32)        */
  double subspace_tol; /* > 0: stop instead when the sine of the largest
                          principal angle between consecutive factors is
                          below it                                         */
  int verbose;         /* per-iteration lines on stdout                    */
} sp_tucker_opts;

void sp_tucker_opts_default(sp_tucker_opts *o);

/* One decomposition run of the host drop-in layer on a context:
 * SpSymtensor::TuckerHOOI / TuckerHOQRI / TuckerAltHOOI
 * (include/detail/sparse/spsymtensor.hpp:40-42, src/sparse/spsymtensor.cpp:
 * 201-337).  U0 (dim x rank, optional) replaces the initial factor.  The
 * reduced-precision forms run their S3TTMc in FP32/BF16 until the error
 * changes by less than switch_tol, then in FP64; a stopping test compares
 * two FP64 iterations only (the first FP64 iteration after the switch never
 * stops the run).
 * Outputs: U_out (dim x rank); per iteration (max_iter entries each, any may
 * be NULL): errors, core_norms (||C||, the packed weighted norm) and times
 * (seconds from the start of the run, initialisation included, to the end
 * of the iteration); the
 * iteration count, the FP64 iterations after the switch and the wall time
 * of the loop. */
int sp_tucker_run(sp_ctx *ctx, const sp_tucker_opts *opts, const double *U0,
                  double *U_out, double *errors, double *core_norms, double *times,
                  uint64_t *iters, uint64_t *fp64_sweeps, double *seconds);

/* Subspace movement of the factor, for convergence tests on the factor
 * rather than on the fit (no counterpart in the reference, whose loops stop
 * on |error change| only, src/sparse/spsymtensor.cpp:235-238).
 * sp_factor_snapshot keeps a device copy of the current U; sp_subspace_sin
 * gives the sine of the largest principal angle between span(U) and the
 * snapshot's span (both orthonormal; resolution about 1e-8). */
int sp_factor_snapshot(sp_ctx *ctx);
int sp_subspace_sin(sp_ctx *ctx, double *out);

/* ---- HOEVD initialisation (replaces hosvd_init's "hoevd" branch,
 * src/sparse/spsymtensor.cpp:363-382, with sparse_syrk 385-409 and
 * convert_to_coo 411-439).
 *
 * sp_hoevd_gram: A = X_(1) X_(1)^T (dim x dim, row-major FP64) of the built
 * tensor, exact for every order and dimension (the reference's base-n column
 * key overflows once dim^(order-1) >= 2^64).
 * sp_hoevd_init: eigenvectors of the `R` largest eigenvalues of A, ascending
 * (as matrix::syevd's last R columns, src/dense/matrixops.cpp:77-100); the
 * result becomes the context's factor and is copied to U (dim x R, row-major)
 * and the eigenvalues to eigvals (R) when those are not NULL.  Up to
 * dim = 4096 A is formed and solved densely (dim^2 below 2^31); above, block
 * Lanczos runs on the implicit A (no dim x dim matrix) when its basis of four
 * blocks of 16*ceil(R/16) columns fits, i.e. about R <= min(dim/4, 512,
 * 1.25e8/dim), and the dense path is used otherwise.  SP_HOEVD=dense|krylov, read at sp_ctx_create,
 * forces one of the two.  SP_ERR_NUMERIC if the Krylov solve does not reach
 * residuals of 1e-12 times the largest eigenvalue. */
int sp_hoevd_gram(sp_ctx *ctx, double *A);
int sp_hoevd_init(sp_ctx *ctx, uint32_t R, double *U, double *eigvals);

/* Eigenvectors of the `rank` largest eigenvalues of the symmetric dim x dim
 * matrix A (host, row-major; only its values are read): U (dim x rank,
 * row-major) with columns in ascending eigenvalue order and w (optional,
 * rank values, ascending) -- matrix::syevd, src/dense/matrixops.cpp:77-100
 * (dsyevd, then columns dim-rank..dim-1), on the context's device: the
 * library's own tridiagonal reduction, bisection and inverse iteration for
 * the `rank` wanted pairs only, and rocSOLVER dormtr (SP_EIG=syevd, syevj or
 * tridiag-rocsolver at context creation select rocSOLVER dsyevd, dsyevj, or
 * rocSOLVER dsytrd for the reduction).  Eigenvector signs are
 * arbitrary, as with LAPACK. */
int sp_syevd_top(sp_ctx *ctx, const double *A, uint64_t dim, uint32_t rank, double *U,
                 double *w);

/* ---- Learned warm start.  No counterpart in the reference: the predicted
 * factor enters at its initialisation hook (hosvd_init,
 * src/sparse/spsymtensor.cpp:339-383) through sp_tucker_run's U0.
 *
 * Node features (dim x 64, FP32) of the context's tensor: log(1+degree),
 * log(1+distinct co-occurrence partners) and a 62-dim sketch of hash-seeded
 * Gaussian node vectors propagated `rounds` times through the owner lists. */
int sp_node_features(sp_ctx *ctx, uint32_t rounds, uint64_t seed, float *out);

/* MLP din -> dh -> dh -> dout (ReLU), FP32, trained with Adam on the mean
 * squared row error; forward, backward and update run in HIP kernels.
 * Parameters are flattened as W1[din][dh], b1, W2[dh][dh], b2, W3[dh][dout], b3. */
typedef struct sp_mlp sp_mlp;
int sp_mlp_create(uint32_t din, uint32_t dh, uint32_t dout, uint64_t seed, sp_mlp **out);
int sp_mlp_destroy(sp_mlp *m);
int sp_mlp_num_params(sp_mlp *m, uint64_t *n);
int sp_mlp_get_params(sp_mlp *m, float *params);
int sp_mlp_set_params(sp_mlp *m, const float *params);
/* Loss and gradient of one batch (host buffers X rows x din, T rows x dout). */
int sp_mlp_loss_grad(sp_mlp *m, const float *X, const float *T, uint64_t rows,
                     double *loss, float *grad);
/* `steps` minibatch Adam steps over seeded row permutations; losses (steps
 * entries) and final_loss may be null. */
int sp_mlp_train(sp_mlp *m, const float *X, const float *T, uint64_t rows,
                 uint32_t steps, uint64_t batch, float lr, uint64_t seed,
                 float *losses, double *final_loss);
int sp_mlp_predict(sp_mlp *m, const float *X, uint64_t rows, float *out);

/* Kernel timing with HIP events on the context's stream.  While enabled,
 * every kernel launch of the library is bracketed by events; sp_timing_get
 * sums the durations of launches whose name matches `name` exactly. */
int sp_timing_enable(sp_ctx *ctx, int on);
int sp_timing_get(sp_ctx *ctx, const char *name, double *total_ms,
                  uint64_t *launches);
int sp_timing_reset(sp_ctx *ctx);

/* Synchronises the context's stream. */
int sp_sync(sp_ctx *ctx);

/* Input side.  .tns reader with the reference's semantics
 * (SymtensLoader, src/utils/loader.cpp:8-36): header "order dim", then per
 * line `order` indices and a value; each tuple is sorted; reading stops at
 * the first failed read.  Call with inds/vals null to get order, dim and
 * nnz, then again with buffers of that size. */
int sp_load_tns(const char *path, uint32_t *order, uint32_t *dim,
                uint64_t *nnz, uint32_t *inds, double *vals);

/* The same reader with a binary tuple cache (no counterpart in the
 * reference).  When `cache` names a file that an earlier call wrote for
 * this source (same size and modification time), the sorted tuples and
 * values are read from it; otherwise the text is parsed and the cache file
 * written (replaced atomically).  Text files are parsed by several host
 * threads; a file whose tokens are not all plain numbers is read by the
 * serial stream reader instead.  The first call of the two-call protocol
 * keeps the tuples in host memory for the second. */
int sp_load_tns_cached(const char *path, const char *cache, uint32_t *order,
                       uint32_t *dim, uint64_t *nnz, uint32_t *inds, double *vals);

/* Seeded synthetic tensors (unique sorted tuples, lexicographically sorted):
 *   kind 0: uniform indices over [1,dim], values 1.0
 *   kind 1: bounded Zipf(s) node popularity over [1,dim], values U(0,1)
 *   kind 2: planted communities (param = p_in, communities = k), values 1.0
 * Draws tuples until nnz unique tuples exist (kind 0 stops after nnz draws
 * and keeps the unique ones).  Returns the count in *nnz_out. */
int sp_generate(int kind, uint32_t order, uint32_t dim, uint64_t nnz,
                double param, uint32_t communities, uint64_t seed,
                uint32_t *inds, double *vals, uint64_t *nnz_out);

#ifdef __cplusplus
}
#endif

#endif /* SYMPROP_AMD_H */
