// Internal declarations of the MI355X S3TTMc library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/symprop_amd.h"
#include "tridiag.hpp"

namespace spamd {

using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;
using u128 = unsigned __int128;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SP_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      throw ::spamd::Error(e_ == hipErrorOutOfMemory ? SP_ERR_OOM           \
                                                     : SP_ERR_DEVICE,       \
                           std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define SP_CHECK_LAUNCH() SP_HIP(hipGetLastError())

// Owning device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    SP_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  // Grows (never shrinks) to at least b bytes; contents are not kept.
  void reserve(size_t b) { if (b > bytes) alloc(b); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Packed-layout tables for one (R, k): lexicographic enumeration of sorted
// k-tuples over [0,R) (symtensor.hpp:129-142) and, for the polynomial
// product p_k = p_{k-1} * (u . x), the gather terms of every output entry.
struct PackedTables {
  u32 R = 0;
  int k = 0;
  u64 P = 0;                      // C(R+k-1, k)
  std::vector<u32> term_ptr;      // P+1
  std::vector<u32> term;          // (src index << 8) | j
  std::vector<double> fact_prod;  // prod_j f_j!   (k-tuple multiplicities)
  std::vector<double> weight;     // k!/prod f_j!  (symtens_coeff)
};
PackedTables make_tables(u32 R, int k);
u64 symsize(u64 R, u64 k);
u64 binom(u64 n, u64 k);

// Device-resident compressed dimension tree.
struct DevLevel {
  u64 size = 0;
  DevBuf keys;       // u64 or u128 packed tuples (key_words x 8 bytes each)
  DevBuf level_idx;  // u32 [size]
  DevBuf level_ptr;  // u64 [size+1]
  DevBuf dep_list;   // u64 [size*(l+1)], l >= 1
};

struct Chunk {
  u64 o_begin, o_end;  // owner range (rem-sorted order)
  u32 r_begin, r_end;  // Y rows written by this chunk
  i64 partial;         // >= 0: piece of one row, result goes to partial slot
  u64 b_begin, b_end;  // batch range (MFMA kernel)
};

// Up to 64 (the fast kernel's batch size) consecutive owners of one Y row; `last` marks the batch after
// which the row is complete (an empty row has one batch with nb = 0).
// All fields are dwords so that the kernel reads a batch with scalar loads.
struct Batch {
  u32 o0_lo, o0_hi;
  u32 row;
  u32 nb_last;  // nb in bits 0..15, last in bit 31
};

struct SplitRow {
  u32 row;
  u32 slot_begin, slot_end;
};

// Work decomposition of the rem-sorted owners for one batch size.
struct Plan {
  std::vector<Chunk> h_chunks;
  std::vector<SplitRow> h_split;
  std::vector<Batch> h_batches;
  DevBuf d_chunks, d_split, d_batches;
  u64 n_partial = 0;
};

struct Timer {
  std::string name;
  hipEvent_t a, b;
};

// A few words of pinned host memory (hipHostMalloc), for device-to-host
// copies that must not wait for the stream: the host reads them after a
// later synchronisation.
struct PinWords {
  int *p = nullptr;
  PinWords() = default;
  PinWords(const PinWords &) = delete;
  PinWords &operator=(const PinWords &) = delete;
  ~PinWords() {
    if (p) (void)hipHostFree(p);
  }
  int *get() {
    if (!p) SP_HIP(hipHostMalloc((void **)&p, 16 * sizeof(int), hipHostMallocDefault));
    return p;
  }
};

// Four timing events, created on first use (hipEventDefault, so elapsed times
// can be read) and destroyed with the owner.
struct EventSet {
  hipEvent_t e[4] = {};
  EventSet() = default;
  EventSet(const EventSet &) = delete;
  EventSet &operator=(const EventSet &) = delete;
  ~EventSet() {
    for (hipEvent_t &x : e)
      if (x) (void)hipEventDestroy(x);
  }
  hipEvent_t *get() {
    for (hipEvent_t &x : e)
      if (!x) SP_HIP(hipEventCreate(&x));
    return e;
  }
};

struct Ctx {
  u32 order = 0, dim = 0;
  u64 nnz = 0;
  int depth = 0;       // order-1 levels
  int bits = 0;        // bits per index in packed keys
  int key_words = 0;   // 1 (u64) or 2 (u128) for the widest level
  double norm = 0.0;
  hipStream_t stream = nullptr;
  int device = 0;

  DevBuf d_inds;  // u32 nnz*order (1-based, sorted rows)
  DevBuf d_vals;  // f64 nnz

  bool built = false;
  bool levels_built = false;  // interior levels and their CSR links
  std::vector<DevLevel> levels;
  DevBuf own_rem;   // u32 [owners], leaf order
  DevBuf own_val;   // f64 [owners], leaf order
  u64 n_owners = 0;

  // S3TTMc work list (owners sorted by rem, then leaf)
  DevBuf w_tuple;   // u32 [owners * m], 0-based rows of the leaf
  DevBuf w_coef;    // f64 [owners], v * multinomial / N!
  DevBuf w_coef32;  // f32 [owners]
  DevBuf w_val;     // f64 [owners], the owner's tensor value
  DevBuf w_leaf;    // u32 [owners], leaf id (index into the last level)
  DevBuf w_segptr;  // u64 [dim+1]
  std::vector<u64> h_segptr;

  // factor / outputs
  u32 R = 0;
  u64 P = 0;        // packed Y width C(R+N-2, N-1)
  DevBuf d_U;       // f64 n x R
  DevBuf d_U32;     // f32 n x R
  DevBuf d_Ubf;     // bf16 n x R
  DevBuf d_Y;       // f64 n x P
  DevBuf d_Y32;     // f32 n x P
  DevBuf d_core;    // f64 R x P
  DevBuf d_W;       // f64 P  (symtens_coeff of order N-1)
  DevBuf d_G;       // f64 P x P
  // rank the device Y and core were computed for (0: none).  A factor of
  // another rank makes them stale: their widths no longer match P.
  u32 y_R = 0;
  u32 core_R = 0;
  DevBuf d_scratch;
  DevBuf d_partial;
  // persistent workspaces of the dense Y-side ops (grown, never shrunk)
  DevBuf ws_rp;    // R x P
  DevBuf ws_pr;    // P x R (top eigenvectors of the Gram)
  DevBuf ws_nr;    // n x R
  DevBuf ws_rr;    // R x R
  DevBuf ws_eig;   // eigen-step vectors (D, E, tau, R top vectors) + info words
  DevBuf ws_tiles; // Gram tile list
  DevBuf ws_uprev; // factor snapshot (n x uprev_R) for subspace angles
  u32 uprev_R = 0;
  int eig_method = 2;  // 2 (default): own tridiagonal reduction (rocSOLVER sytrd
                       // beyond its LDS limit) + host top-R tridiagonal solve +
                       // rocSOLVER ormtr; 3: as 2 with rocSOLVER sytrd always;
                       // 0: rocSOLVER syevd (divide and conquer); 1: syevj
  int hoevd_method = 0;  // HOEVD solver: 0 (default) dense for dim <= 4096, block
                         // Krylov on the implicit operator above; 1: dense; 2: Krylov
  DevBuf ws_sym;   // symmetric matrix uploaded by sp_syevd_top
  DevBuf ws_trd;   // tridiagonal reduction: broadcast vectors, barrier words
  DevBuf ws_trd_v; // tridiagonal reduction: reflector tails (P x P)
  DevBuf ws_amax;  // eigen-step: per-block maxima of |G|, then the scaling exponent
  PinWords h_pin;  // [0]: leftsvd's scaling exponent (scale_to_unit_async)
  // polls of one exchange wait before the reduction gives up (about a second;
  // SP_TRD_SPIN at context creation sets it, to test the fallback)
  unsigned trd_spin_limit = 1u << 21;
  // tridiagonal reduction: 0 (default) one workgroup for P <= 256, rows in
  // LDS across workgroups up to P = 1,024, rows in registers across
  // workgroups above; 1 (SP_SYTRD=wg at context creation) rows in LDS for
  // every P; 2 (SP_SYTRD=reg) rows in registers for every P
  int trd_mode = 0;
  u64 trd_timeouts = 0;  // reductions that timed out and fell back
  // leftsvd's warm-started Rayleigh-Ritz eigen-step (dense.hip): on by default
  // (SP_EIG_WARM=0 at context creation turns it off).  After a miss the next
  // warm_skip calls go straight to the direct solver, sized from the measured
  // times (running means, ms, taken with eig_ev on the stream so that work
  // queued before the step is not counted) so that misses cost at most a
  // twentieth of the direct path's time; a near miss or a success clears it
  bool warm_enabled = true;
  u32 warm_skip = 0;
  u32 kry_m = 0;  // blocks of the Krylov form for P <= 256 (0: not chosen yet)
  u32 kry_floor = 2, kry_hold = 0;  // least m while kry_hold calls remain
  double warm_miss_ms = 0.0, direct_eig_ms = 0.0;
  EventSet eig_ev;
  // host staging of the dense Y-side steps, kept across calls so that a
  // host-to-device copy from them needs no wait before the function returns
  // (the next write to one follows a synchronisation of the stream)
  std::vector<double> h_eig_de, h_eig_z, h_theta, h_small;
  int h_trd_abort = 0;
  u64 gram_tiles_P = 0;  // P of the Gram tile list in ws_tiles (0: none)
  DevBuf ws_warm;  // its bases, products and small matrices

  // chunking of the work list for the current R: plan[0] with 64-owner
  // batches (generic, FP32 and BF16 kernels), plan[1] with the FP64 fast
  // kernel's batch size when that is smaller
  Plan plan[2];
  u32 chunk_R = 0;

  // tables per level for the current R
  std::vector<PackedTables> tabs;  // index k = 1..m
  DevBuf d_tab_ptr, d_tab_term, d_fact;
  std::vector<u64> tab_ptr_off, tab_term_off;

  // compile-time specialised MFMA kernel for (order-1, R), if instantiated
  void *fast_kernel = nullptr;
  void *fast_kernel_f32 = nullptr;
  void *fast_kernel_bf16 = nullptr;
  int fast_H = 0;
  int fast_bo = 64;  // owners per batch of the fast kernels
  DevBuf d_fold_ptr, d_fold_term;  // fold lists of the MFMA kernels (CSR)
  DevBuf d_fold_tab;               // the same, thread-major
  int fold_len = 0;

  // profiling
  bool profile = false;
  std::vector<Timer> timers;

  std::string last_error;
};

// construction (cdtree.hip)
// Checks the uploaded tuples (indices in [1, dim], sorted) and sets c.norm.
void input_check_norm(Ctx &c);
void cdtree_build(Ctx &c);
void cdtree_ensure_levels(Ctx &c);
void cdtree_ensure_dep_list(Ctx &c, int l);
void cdtree_export_level(Ctx &c, int l, u32 *keys, u32 *level_idx,
                         u64 *level_ptr, u64 *dep_list);
void cdtree_export_owners(Ctx &c, u32 *rem, double *val);

// sort / scan primitives (sort.hip)
void exclusive_scan_u64(const u64 *in, u64 *out, u64 n, hipStream_t s);
template <class K>
void radix_sort(K *keys, K *keys_tmp, u32 *vals, u32 *vals_tmp, u64 n,
                int bits, hipStream_t s, K **keys_out, u32 **vals_out);

// S3TTMc (s3ttmc.hip, s3ttmc_fast.hip)
void s3ttmc_prepare(Ctx &c, u32 R);
void s3ttmc_run(Ctx &c, int precision);
// Queues the S3TTMc kernels of the prepared rank (c.chunk_R) on c.stream:
// reads the FP64 factor U (dim x R), writes its FP32 copy to U32 for the
// reduced precisions, the packed result to Y (dim x P) and split-row partial
// sums to `partial` (s3ttmc_partial_slots x P doubles).  `timed`: bracket the
// launches with the context's profiling events when profiling is on.
void s3ttmc_launch(Ctx &c, int precision, const double *U, float *U32, double *Y,
                   double *partial, bool timed);
u64 s3ttmc_partial_slots(const Ctx &c, int precision);
// Allocation-free form of the launch above (noalloc.hip).
void noalloc_setup(int device);
void s3ttmc_noalloc(Ctx &c, const double *U, u32 R, double *Y, int precision);
void make_fold_table(u32 R, int M, int H, std::vector<u32> &ptr, std::vector<u32> &term,
                     std::vector<u32> &tab, int &len);
void *fast_kernel_for(int M, u32 R, int *H, int *bo, void **kf32, void **kbf16);
void launch_fast(void *kernel, unsigned grid, hipStream_t s, const Chunk *chunks,
                 const Batch *batches, const u32 *tup, const double *coef, const double *U,
                 const float *coef32, const float *U32, const u32 *fold_ptr,
                 const u32 *fold_term, const u32 *fold_tab, int fold_len, const double *fact,
                 double *Y, double *partial);

// dense Y-side ops (dense.hip)
void dense_gemmtn(Ctx &c);
void dense_symgemm(Ctx &c);
void dense_gram(Ctx &c);
void dense_leftsvd_y(Ctx &c, double *sigma);
void dense_qr(Ctx &c);
void dense_regularize(Ctx &c, double delta);
double dense_core_norm(Ctx &c);
void dense_factor_snapshot(Ctx &c);
// out = A^T B (A: n x R, B: n x P, row-major; out R x P), fixed summation order.
void dense_atb(Ctx &c, const double *A, u32 R, const double *B, u64 P, u64 n, double *out);
// out = Y M^T (Y: n x P, M: R x P, out: n x R), queued on the context's stream.
void dense_ymt(Ctx &c, const double *Y, u64 n, u64 P, const double *M, u32 R, double *out);
double dense_subspace_sin(Ctx &c);
// Host: eigen-decomposition of a small symmetric matrix (row-major n x n),
// eigenvalues descending in w, eigenvectors in the columns of V (row-major).
void dense_small_eig(const std::vector<double> &A, int n, std::vector<double> &w,
                     std::vector<double> &V);

// HOEVD initial factor (hoevd.hip)
void hoevd_gram(Ctx &c, double *A);
void hoevd_init(Ctx &c, u32 R, double *U_out, double *eig_out);

// Scales the device array A (count values) by 2^ex so that its largest |a|
// lies in [1, 2) when that largest value is outside [2^-400, 2^400] (finite,
// nonzero); returns ex (0: not scaled).
int scale_to_unit(Ctx &c, double *A, u64 count);
// The same decision and scaling queued on the stream without a wait: the
// exponent lands in c.h_pin.get()[0] once the stream has passed this point
// (read it only after a later synchronisation of the stream).
void scale_to_unit_async(Ctx &c, double *A, u64 count);

// Top-R eigenpairs of the P x P symmetric matrix G (device, destroyed) by
// rocSOLVER (c.eig_method): V (device, P x R row-major) and theta (host),
// descending.
void eig_top(Ctx &c, double *G, u64 P, u32 R, double *V, std::vector<double> &theta,
             bool in_range = false);

// Tridiagonal reduction of the symmetric P x P G (device, only read) in one
// launch of co-resident workgroups (sytrd.hip): D, E, tau and the reflector
// tails in Vr (P x P) in LAPACK's uplo = lower layout, queued on the
// context's stream.  Returns false, with nothing launched, when P does not
// fit the kernel's LDS budget; otherwise *abort_flag points to a device word
// that is nonzero once the kernel has run if a wait between its workgroups
// timed out (then every workgroup left, G is untouched and D, E, tau, Vr are
// not valid), or is null for the one-workgroup form, which has no such wait.
bool sytrd_lower(Ctx &c, const double *G, u64 P, double *D, double *E, double *tau, double *Vr,
                 const int **abort_flag);
constexpr u64 kTrdMaxP = 2048;  // sytrd_lower returns false beyond this order

// Orthonormal rows Qt (m b x P) of the block Krylov space of the b rows X0t
// (b x P) under the symmetric G (P x P), and GQt = Qt G, in one workgroup
// (eig_krylov.hip); rows past a breakdown (an invariant subspace) are zero.
// Queued on the context's stream; false, with nothing launched, unless
// P <= 256, b <= 16 and m b <= P.
bool krylov_basis(Ctx &c, const double *G, const double *X0t, u64 P, u32 b, u32 m, double *Qt,
                  double *GQt);

// V = Q Z (V row-major P x R, Z column-major P x R, both device) for the Q of
// sytrd_lower's output in Vr (passed as G) and tau (sytrd.hip); queued on the
// context's stream.  Returns false, with nothing done, for P outside 2..2048.
bool apply_q_lower(Ctx &c, const double *G, const double *tau, u64 P, const double *Z, u32 R,
                   double *V);

// learned warm start (learned.hip)
struct Mlp;
Mlp *mlp_create(int din, int dh, int dout, u64 seed);
void mlp_destroy(Mlp *m);
long mlp_num_params(Mlp *m);
void mlp_get_params(Mlp *m, float *out);
void mlp_set_params(Mlp *m, const float *in);
double mlp_loss_grad(Mlp *m, const float *X, const float *T, long rows, float *grad);
double mlp_train(Mlp *m, const float *X, const float *T, long rows, int steps, long batch,
                 float lr, u64 seed, float *losses);
void mlp_predict(Mlp *m, const float *X, long rows, float *out);
void node_features(Ctx &c, int rounds, u64 seed, float *f);

// error string of the C-ABI (sp_api.hip)
void set_last_error(const std::string &msg);

// input side (input.cpp)
double host_norm(u32 order, u64 nnz, const u32 *inds, const double *vals);

// timing helpers
struct ScopedTimer {
  Ctx &c;
  Timer t;
  bool on;
  ScopedTimer(Ctx &cx, const char *name) : c(cx), on(cx.profile) {
    if (!on) return;
    t.name = name;
    SP_HIP(hipEventCreate(&t.a));
    SP_HIP(hipEventCreate(&t.b));
    SP_HIP(hipEventRecord(t.a, c.stream));
  }
  ~ScopedTimer() {
    if (!on) return;
    (void)hipEventRecord(t.b, c.stream);
    c.timers.push_back(t);
  }
};

// Kernel attributes that the launchers need, set or queried once per
// (device, kernel) per process instead of at every launch: the dynamic-LDS
// limit is raised only when a launch asks for more than was allowed so far,
// and the static LDS size is read once.
inline void allow_dyn_lds(int device, const void *kern, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> allowed;
  std::lock_guard<std::mutex> lock(mu);
  size_t &cur = allowed[{device, kern}];
  if (bytes <= cur) return;
  SP_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  cur = bytes;
}

inline size_t static_lds(int device, const void *kern) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> known;
  std::lock_guard<std::mutex> lock(mu);
  auto it = known.find({device, kern});
  if (it != known.end()) return it->second;
  hipFuncAttributes fa;
  SP_HIP(hipFuncGetAttributes(&fa, kern));
  known[{device, kern}] = fa.sharedSizeBytes;
  return fa.sharedSizeBytes;
}

inline unsigned grid_for(u64 n, unsigned block) {
  u64 g = (n + block - 1) / block;
  return (unsigned)(g == 0 ? 1 : g);
}

}  // namespace spamd
