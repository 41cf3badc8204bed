// Host C++ implementation of the drop-in decomposition API (see symtucker.hpp).
#include "symtucker.hpp"

#include "../../include/symprop_amd.h"
#include "sp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <random>
#include <stdexcept>
#include <map>

namespace symprop {

namespace {

void ck(int rc) {
  if (rc != SP_OK)
    throw std::runtime_error(std::string("symprop_amd: ") + sp_last_error());
}

index_t binom(index_t n, index_t k) {
  if (k > n) return 0;
  if (k * 2 > n) k = n - k;
  if (k == 0) return 1;
  index_t r = n;
  for (index_t i = 2; i <= k; ++i) {
    r *= (n - i + 1);
    r /= i;
  }
  return r;
}

// spsymtensor.cpp:148-158
real_t rel_residual(const std::vector<real_t> &x, const std::vector<real_t> &y) {
  real_t residual = 0.0, norm = 0.0;
  for (size_t i = 0; i < x.size(); i++) {
    residual += std::pow((x[i] - y[i]), 2);
    norm += std::pow(x[i], 2);
  }
  return residual / norm;
}

// symtensor.cpp:78-106
real_t symtens_rel_res(const std::vector<real_t> &c1, const std::vector<real_t> &c2,
                       const std::vector<real_t> &coeff, size_t nrows) {
  real_t residual = 0.0, norm = 0.0;
  const size_t P = coeff.size();
  for (size_t i = 0; i < nrows; i++)
    for (size_t j = 0; j < P; j++) {
      residual += std::pow(c1[i * P + j] - c2[i * P + j], 2) * coeff[j];
      norm += std::pow(c1[i * P + j], 2) * coeff[j];
    }
  return residual / norm;
}

real_t symtens_norm(const std::vector<real_t> &core, const std::vector<real_t> &coeff,
                    size_t nrows) {
  real_t norm = 0.0;
  const size_t P = coeff.size();
  for (size_t i = 0; i < nrows; i++)
    for (size_t j = 0; j < P; j++) norm += std::pow(core[i * P + j], 2) * coeff[j];
  return std::sqrt(norm);
}

real_t plain_norm(const std::vector<real_t> &x) {
  real_t n = 0.0;
  for (real_t v : x) n += std::pow(v, 2);
  return std::sqrt(n);
}

// sqrt(||X||^2 - ||C||^2) / ||X|| (spsymtensor.cpp:230-238, tucker.cpp:40-45).
// For an exactly fitted tensor the difference is a rounding residue of either
// sign; the reference's sqrt of a negative residue is NaN, which no stopping
// rule accepts.  It is clamped to 0 here.
real_t fit_error(real_t norm, real_t core_norm) {
  return std::sqrt(std::max(norm * norm - core_norm * core_norm, real_t(0))) / norm;
}

real_t auto_switch_tol(const Config &c) {
  if (c.switch_tol > 0) return c.switch_tol;
  const real_t t = (c.precision == SP_PREC_BF16) ? 1e-3 : 1e-5;
  return std::max(t, c.rel_tol);
}

// Precision schedule and stopping test of the decomposition loops.  The
// reference stops when the error differs from the previous iteration's by
// less than rel_tol (spsymtensor.cpp:235-238).  The reduced-precision forms
// run S3TTMc in FP32/BF16 until the error changes by less than the switch
// tolerance, then in FP64; the test only ever compares two FP64 iterations,
// so the first FP64 iteration after the switch never stops the run.  With
// subspace_tol > 0 the test is on the factor instead: the sine of the
// largest principal angle between consecutive factors.
class IterationControl {
public:
  IterationControl(const Config &c, sp_ctx *ctx)
      : _ctx(ctx), _prec(c.precision), _low(c.precision != SP_PREC_FP64),
        _switch_tol(auto_switch_tol(c)), _rel_tol(c.rel_tol), _sub_tol(c.subspace_tol) {}

  int precision() const { return _low ? _prec : SP_PREC_FP64; }
  index_t fp64_sweeps() const { return _fp64; }

  // Called once the iteration's factor is final; returns true to stop.
  bool finish(real_t error_diff) {
    real_t moved = 1.0;
    if (_sub_tol > 0) {
      if (_have_snapshot) ck(sp_subspace_sin(_ctx, &moved));
      ck(sp_factor_snapshot(_ctx));
      _have_snapshot = true;
    }
    if (_low) {
      if (error_diff < _switch_tol) _low = false;
      _prev_fp64 = false;
      return false;
    }
    if (_prec != SP_PREC_FP64) ++_fp64;
    const bool comparable = _prec == SP_PREC_FP64 || _prev_fp64;
    _prev_fp64 = true;
    if (!comparable) return false;
    return _sub_tol > 0 ? moved < _sub_tol : error_diff < _rel_tol;
  }

private:
  sp_ctx *_ctx;
  int _prec;
  bool _low;
  real_t _switch_tol, _rel_tol, _sub_tol;
  index_t _fp64 = 0;
  bool _prev_fp64 = false;
  bool _have_snapshot = false;
};

}  // namespace

template <class F>
void SpSymtensor::timed(const char *name, F &&f) const {
  if (!_phase_timing) {
    f();
    return;
  }
  FunctionTimer t(name);
  f();
  ck(sp_sync(_ctx));
}

// ------------------------------------------------------------ packed helpers
index_t symtens_size(index_t dim, index_t order) { return binom(order + dim - 1, order); }

std::vector<real_t> symtens_coeff(size_t dim, size_t order) {
  std::vector<real_t> out;
  std::vector<dim_t> idx(order, 0);
  auto fact = [](int n) {
    real_t f = 1;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
  };
  while (true) {
    real_t d = 1;
    int run = 1;
    for (size_t q = 1; q <= order; ++q) {
      if (q < order && idx[q] == idx[q - 1]) ++run;
      else { d *= fact(run); run = 1; }
    }
    out.push_back(fact((int)order) / d);
    int p = (int)order - 1;
    while (p >= 0 && idx[p] == dim - 1) --p;
    if (p < 0) break;
    idx[p]++;
    for (size_t q = p + 1; q < order; ++q) idx[q] = idx[p];
  }
  return out;
}

// symtensor.hpp:96-116: every distinct permutation of each packed entry
void unpack_symtens(const real_t *packed, dim_t dim, order_t order, real_t *full) {
  std::vector<dim_t> idx(order, 0);
  index_t p = 0;
  while (true) {
    std::vector<dim_t> perm = idx;
    do {
      index_t lin = 0;
      for (order_t j = 0; j < order; ++j) lin = lin * dim + perm[j];
      full[lin] = packed[p];
    } while (std::next_permutation(perm.begin(), perm.end()));
    ++p;
    int q = (int)order - 1;
    while (q >= 0 && idx[q] == dim - 1) --q;
    if (q < 0) break;
    idx[q]++;
    for (order_t r = q + 1; r < order; ++r) idx[r] = idx[q];
  }
}

// ------------------------------------------------------------ FunctionTimer
namespace {
// Per-name accumulated wall time (microseconds) and call count.
struct TimerTotals {
  long long micros = 0;
  long long calls = 0;
};
std::mutex g_timer_mutex;
std::map<std::string, TimerTotals> &timer_table() {
  static std::map<std::string, TimerTotals> table;
  return table;
}
}  // namespace

FunctionTimer::FunctionTimer(std::string_view name)
    : m_start(std::chrono::high_resolution_clock::now()), m_name(name) {}
FunctionTimer::~FunctionTimer() { stop(); }

void FunctionTimer::stop() {
  if (m_stopped) return;
  m_stopped = true;
  const long long us = std::chrono::duration_cast<std::chrono::microseconds>(
                           std::chrono::high_resolution_clock::now() - m_start)
                           .count();
  std::lock_guard<std::mutex> lk(g_timer_mutex);
  TimerTotals &t = timer_table()[m_name];
  t.micros += us;
  t.calls += 1;
}

// Report layout of the reference (src/utils/function_timer.cpp:25-46): a
// header line, then one line per name, longest total first, as
// "%10.3f ms  | %10lld calls | name" with the total in milliseconds.
void FunctionTimer::printTimes() {
  std::vector<std::pair<std::string, TimerTotals>> rows;
  {
    std::lock_guard<std::mutex> lk(g_timer_mutex);
    rows.assign(timer_table().begin(), timer_table().end());
  }
  std::stable_sort(rows.begin(), rows.end(), [](const auto &x, const auto &y) {
    return x.second.micros > y.second.micros;
  });
  std::string text = "Accumulated function times:\n";
  char line[64];
  for (const auto &[name, t] : rows) {
    std::snprintf(line, sizeof line, "%10.3f ms  | %10lld calls | ", t.micros / 1000.0,
                  t.calls);
    text += line;
    text += name;
    text += '\n';
  }
  std::cout << text;
}

void FunctionTimer::clear() {
  std::lock_guard<std::mutex> lk(g_timer_mutex);
  timer_table().clear();
}

// ------------------------------------------------------------ SymTuckerDecomp
// tucker.cpp:25-36
SymTuckerDecomp::SymTuckerDecomp(const Config &config, dim_t dim, order_t order)
    : _dim(dim), _rank(config.rank), _order(order) {
  _core.resize((size_t)std::pow(_rank, order));
  _factor.resize((size_t)dim * _rank);
  std::normal_distribution<real_t> distr(0.0, 1.0);
  std::mt19937 engine{config.seed};
  for (size_t i = 0; i < _factor.size(); i++) _factor[i] = distr(engine);
}

// tucker.cpp:40-45
real_t SymTuckerDecomp::reconstruct_error(const SpSymtensor &target) const {
  return fit_error(target.norm(), plain_norm(_core));
}

// tucker.cpp:47-61, including its core indexing _core[i * _rank + j]
void SymTuckerDecomp::print(std::ostream &os) const {
  os << "Mode-1 matricized C:\n";
  for (size_t i = 0; i < _rank; i++) {
    for (size_t j = 0; j < _core.size() / _rank; j++) os << _core[i * _rank + j] << ",";
    os << "\n";
  }
  os << "Factor U:\n";
  for (size_t i = 0; i < _dim; i++) {
    for (size_t j = 0; j < _rank; j++) os << _factor[i * _rank + j] << ",";
    os << "\n";
  }
}

// ------------------------------------------------------------ SpSymtensor
SpSymtensor::SpSymtensor(dim_t dim, order_t order, const std::vector<dim_t> &unnz_inds,
                         const std::vector<real_t> &unnzs)
    : _dim(dim), _order(order), _unnz(unnzs.size()), _owns(true) {
  START_FUNCTION_TIMER();
  if (unnz_inds.size() != (size_t)_unnz * order)
    throw std::invalid_argument("index array size does not match nnz * order");
  ck(sp_ctx_create(order, dim, _unnz, unnz_inds.data(), unnzs.data(), &_ctx));
  ck(sp_norm(_ctx, &_norm));
  {
    FunctionTimer t("CompressedDTree");
    ck(sp_cdtree_build(_ctx));
  }
  std::cout << "Finished creating compressed dtree\n";
}

SpSymtensor::SpSymtensor(sp_ctx *ctx) : _ctx(ctx), _owns(false) {
  ck(sp_ctx_info(_ctx, &_order, &_dim, &_unnz));
  ck(sp_norm(_ctx, &_norm));
}

SpSymtensor::~SpSymtensor() {
  if (_owns && _ctx) sp_ctx_destroy(_ctx);
}

// spsymtensor.cpp:30-110 (packed result; unpacked rows when !partial_sym_res)
void SpSymtensor::S3TTMc(std::vector<real_t> &res, const std::vector<real_t> &U, dim_t R,
                         bool partial_sym_res) const {
  START_FUNCTION_TIMER();
  const size_t P = symtens_size(R, _order - 1);
  const size_t full = (size_t)std::pow(R, _order - 1);
  const size_t ncols = partial_sym_res ? P : full;
  if (res.size() != (size_t)_dim * ncols) res.resize((size_t)_dim * ncols);
  if (U.size() != (size_t)_dim * R) throw std::invalid_argument("U must be dim x R");
  ck(sp_factor_upload(_ctx, U.data(), R));
  ck(sp_s3ttmc_dev(_ctx, SP_PREC_FP64));
  ck(sp_y_download(_ctx, res.data(), ncols));
  if (!partial_sym_res) {
    std::vector<real_t> row(P);
    for (size_t i = 0; i < _dim; i++) {
      std::copy(res.begin() + i * ncols, res.begin() + i * ncols + P, row.begin());
      unpack_symtens(row.data(), R, _order - 1, res.data() + i * ncols);
    }
  }
}

// spsymtensor.cpp:339-383
void SpSymtensor::init_factor(std::vector<real_t> &U, const Config &c) const {
  FunctionTimer t("hosvd_init");
  if (c.init == "hosvd") {
    printf("Error: sparse SVD requires MKL.\n");
    return;
  }
  if (c.init == "hoevd") {
    hoevd_init(U, c);
    return;
  }
  if (c.init == "given") {
    if (c.U0.size() != U.size()) throw std::invalid_argument("U0 must be dim x rank");
    U = c.U0;
  }
}

// spsymtensor.cpp:363-382.  The reference also prints the whole dim x dim
// matrix A to stdout (374-379); that debug dump is reproduced only when the
// environment sets SP_HOEVD_PRINT=1.
void SpSymtensor::hoevd_init(std::vector<real_t> &U, const Config &c) const {
  const char *pr = std::getenv("SP_HOEVD_PRINT");
  if (pr && std::string(pr) == "1") {
    std::vector<real_t> A((size_t)_dim * _dim);
    ck(sp_hoevd_gram(_ctx, A.data()));
    for (size_t i = 0; i < _dim; i++) {
      for (size_t j = 0; j < _dim; j++) std::cout << A[i * _dim + j] << ",";
      std::cout << "\n";
    }
  }
  U.resize((size_t)_dim * c.rank);
  ck(sp_hoevd_init(_ctx, c.rank, U.data(), nullptr));
}

// spsymtensor.cpp:201-252
SymTuckerDecomp SpSymtensor::TuckerHOOI(const Config &c) const {
  START_FUNCTION_TIMER();
  const auto t0 = std::chrono::steady_clock::now();
  SymTuckerDecomp decomp(c, _dim, _order);
  auto &core = decomp._core;
  auto &factor = decomp._factor;
  init_factor(factor, c);
  const dim_t R = c.rank;
  const size_t P = symtens_size(R, _order - 1);
  const size_t full = (size_t)std::pow(R, _order - 1);
  // the core stays packed inside the loop: the unpacked core's residual and
  // norm are sums over the packed entries weighted by their multiplicities
  // (R x P terms instead of R^N); it is unpacked once at the end
  std::vector<real_t> core_p((size_t)R * P), core_prev((size_t)R * P, 0.0);
  const std::vector<real_t> coeff = symtens_coeff(R, _order - 1);
  ck(sp_factor_upload(_ctx, factor.data(), R));
  IterationControl ctl(c, _ctx);
  real_t prev_error = 0.0;
  for (size_t iter = 0; iter < c.max_iter; iter++) {
    timed("regularize", [&] { ck(sp_regularize(_ctx, c.delta)); });
    timed("S3TTMc", [&] { ck(sp_s3ttmc_dev(_ctx, ctl.precision())); });
    timed("leftsvd", [&] { ck(sp_leftsvd_y(_ctx, nullptr)); });
    timed("gemmtn", [&] { ck(sp_gemmtn(_ctx, core_p.data())); });
    const real_t rel_res = symtens_rel_res(core_prev, core_p, coeff, R);
    const real_t core_norm = symtens_norm(core_p, coeff, R);
    const real_t error = fit_error(_norm, core_norm);  // reconstruct_error
    if (c.verbose)
      std::cout << std::scientific << std::setprecision(12) << "Iteration: " << iter
                << " relative residual: " << rel_res << " error " << error << "\n";
    decomp.record(error, core_norm, t0);
    const real_t error_diff = std::abs(prev_error - error);
    if (ctl.finish(error_diff)) break;
    prev_error = error;
    core_prev = core_p;
  }
  for (dim_t r = 0; r < R; ++r)
    unpack_symtens(core_p.data() + (size_t)r * P, R, _order - 1, core.data() + r * full);
  ck(sp_factor_download(_ctx, factor.data()));
  decomp._fp64_sweeps = ctl.fp64_sweeps();
  decomp._seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return decomp;
}

// spsymtensor.cpp:254-299
SymTuckerDecomp SpSymtensor::TuckerHOQRI(const Config &c) const {
  START_FUNCTION_TIMER();
  const auto t0 = std::chrono::steady_clock::now();
  SymTuckerDecomp decomp(c, _dim, _order);
  auto &factor = decomp._factor;
  init_factor(factor, c);
  const dim_t R = c.rank;
  const size_t ncols = symtens_size(R, _order - 1);
  std::vector<real_t> core((size_t)R * ncols), core_prev((size_t)R * ncols);
  const std::vector<real_t> coeff = symtens_coeff(R, _order - 1);
  ck(sp_factor_upload(_ctx, factor.data(), R));
  IterationControl ctl(c, _ctx);
  real_t prev_error = 0.0;
  for (size_t iter = 0; iter < c.max_iter; iter++) {
    timed("regularize", [&] { ck(sp_regularize(_ctx, c.delta)); });
    timed("S3TTMc", [&] { ck(sp_s3ttmc_dev(_ctx, ctl.precision())); });
    timed("gemmtn", [&] { ck(sp_gemmtn(_ctx, core.data())); });
    timed("symgemm", [&] { ck(sp_symgemm(_ctx)); });
    timed("qr", [&] { ck(sp_qr(_ctx)); });
    const real_t rel_res = symtens_rel_res(core, core_prev, coeff, R);
    const real_t core_norm = symtens_norm(core, coeff, R);
    const real_t error = fit_error(_norm, core_norm);
    if (c.verbose)
      std::cout << std::scientific << std::setprecision(12) << "Iteration: " << iter
                << " relative residual: " << rel_res << " error " << error << "\n";
    decomp.record(error, core_norm, t0);
    core_prev = core;
    const real_t error_diff = std::abs(prev_error - error);
    if (ctl.finish(error_diff)) break;
    prev_error = error;
  }
  const size_t unpacked = (size_t)std::pow(R, _order - 1);
  for (dim_t i = 0; i < R; i++)
    unpack_symtens(core.data() + (size_t)i * ncols, R, _order - 1,
                   decomp._core.data() + (size_t)i * unpacked);
  ck(sp_factor_download(_ctx, factor.data()));
  decomp._fp64_sweeps = ctl.fp64_sweeps();
  decomp._seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return decomp;
}

// spsymtensor.cpp:301-337.  The n x n symsyrk + syevd pair is replaced by the
// P x P weighted Gram of the packed Y (same leading eigenvectors); columns
// are returned in ascending eigenvalue order as matrixops.cpp:91-97 does.
SymTuckerDecomp SpSymtensor::TuckerAltHOOI(const Config &c) const {
  START_FUNCTION_TIMER();
  const auto t0 = std::chrono::steady_clock::now();
  SymTuckerDecomp decomp(c, _dim, _order);
  auto &core = decomp._core;
  auto &factor = decomp._factor;
  auto factor_prev = factor;
  const std::vector<real_t> coeff = symtens_coeff(c.rank, _order - 1);
  init_factor(factor, c);
  const dim_t R = c.rank;
  const size_t ncols = symtens_size(R, _order - 1);
  std::vector<real_t> core_p((size_t)R * ncols);
  ck(sp_factor_upload(_ctx, factor.data(), R));
  IterationControl ctl(c, _ctx);
  real_t prev_error = 0.0;
  for (size_t iter = 0; iter < c.max_iter; iter++) {
    timed("regularize", [&] { ck(sp_regularize(_ctx, c.delta)); });
    timed("S3TTMc", [&] { ck(sp_s3ttmc_dev(_ctx, ctl.precision())); });
    timed("syevd", [&] {
      ck(sp_leftsvd_y(_ctx, nullptr));
      ck(sp_factor_download(_ctx, factor.data()));
      for (size_t i = 0; i < _dim; ++i)
        std::reverse(factor.begin() + i * R, factor.begin() + (i + 1) * R);
      ck(sp_factor_upload(_ctx, factor.data(), R));
    });
    const real_t rel_res = rel_residual(factor, factor_prev);
    timed("gemmtn", [&] { ck(sp_gemmtn(_ctx, core_p.data())); });
    std::copy(core_p.begin(), core_p.end(), core.begin());
    const real_t core_norm = symtens_norm(core_p, coeff, R);
    const real_t error = fit_error(_norm, core_norm);
    if (c.verbose)
      std::cout << std::scientific << std::setprecision(10) << "Iteration: " << iter
                << " relative residual: " << rel_res << " error " << error << "\n";
    decomp.record(error, core_norm, t0);
    const real_t error_diff = std::abs(prev_error - error);
    if (ctl.finish(error_diff)) break;
    prev_error = error;
    factor_prev = factor;
  }
  decomp._fp64_sweeps = ctl.fp64_sweeps();
  decomp._seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return decomp;
}

}  // namespace symprop

// ------------------------------------------------------------ C entry
extern "C" void sp_tucker_opts_default(sp_tucker_opts *o) {
  if (!o) return;
  const symprop::Config c;
  o->alg = 0;
  o->rank = c.rank;
  o->max_iter = c.max_iter;
  o->rel_tol = c.rel_tol;
  o->delta = c.delta;
  o->seed = c.seed;
  o->init = "rnd";
  o->precision = c.precision;
  o->switch_tol = c.switch_tol;
  o->subspace_tol = c.subspace_tol;
  o->verbose = 0;
}

extern "C" int sp_tucker_run(sp_ctx *ctx, const sp_tucker_opts *o, const double *U0,
                             double *U_out, double *errors, double *core_norms, double *times,
                             uint64_t *iters, uint64_t *fp64_sweeps, double *seconds) {
  using namespace symprop;
  try {
    if (!ctx || !o || !U_out || !iters) throw spamd::Error(SP_ERR_ARG, "null argument");
    if (o->alg < 0 || o->alg > 2) throw spamd::Error(SP_ERR_ARG, "alg must be 0, 1 or 2");
    if (o->rank < 1) throw spamd::Error(SP_ERR_ARG, "rank must be >= 1");
    if (o->precision < SP_PREC_FP64 || o->precision > SP_PREC_BF16)
      throw spamd::Error(SP_ERR_ARG, "unknown precision");
    SpSymtensor T(ctx);
    // iteration times are recorded by the loop itself; no per-phase timers
    // and no waits between the phases of an iteration
    T.set_phase_timing(false);
    Config c;
    c.rank = o->rank;
    c.max_iter = o->max_iter;
    c.rel_tol = o->rel_tol;
    c.delta = o->delta;
    c.seed = o->seed;
    c.init = o->init ? o->init : "rnd";
    c.precision = o->precision;
    c.switch_tol = o->switch_tol;
    c.subspace_tol = o->subspace_tol;
    c.verbose = o->verbose != 0;
    if (U0) {
      c.init = "given";
      c.U0.assign(U0, U0 + (size_t)T.dim() * o->rank);
    }
    const SymTuckerDecomp r = o->alg == 0   ? T.TuckerHOOI(c)
                              : o->alg == 1 ? T.TuckerHOQRI(c)
                                            : T.TuckerAltHOOI(c);
    std::copy(r.factor().begin(), r.factor().end(), U_out);
    if (errors) std::copy(r.errors().begin(), r.errors().end(), errors);
    if (core_norms) std::copy(r.core_norms().begin(), r.core_norms().end(), core_norms);
    if (times) std::copy(r.times().begin(), r.times().end(), times);
    *iters = r.iterations();
    if (fp64_sweeps) *fp64_sweeps = r.fp64_sweeps();
    if (seconds) *seconds = r.seconds();
    return SP_OK;
  } catch (const spamd::Error &e) {
    spamd::set_last_error(e.what());
    return e.code;
  } catch (const std::invalid_argument &e) {
    spamd::set_last_error(e.what());
    return SP_ERR_ARG;
  } catch (const std::exception &e) {
    spamd::set_last_error(e.what());
    return SP_ERR_DEVICE;
  }
}
