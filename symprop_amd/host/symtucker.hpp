// Host C++ drop-in for SymProp's decomposition API on the MI355X S3TTMc path.
//
// Same names and argument meaning as the reference:
//   Config            include/detail/tucker.hpp:14-23
//   SymTuckerDecomp   include/detail/tucker.hpp:25-45, src/tucker.cpp:25-61
//   SpSymtensor       include/detail/sparse/spsymtensor.hpp:26-65
//   FunctionTimer     include/utils/function_timer.hpp, src/utils/function_timer.cpp
// Everything heavy runs on the GPU through the C-ABI of include/symprop_amd.h;
// the tensor, its tree and the Y / core / factor iterates stay in HBM.
#pragma once

#include <chrono>
#include <cstdint>
#include <iosfwd>
#include <string>
#include <string_view>
#include <vector>

struct sp_ctx;

namespace symprop {

using real_t = double;
using order_t = uint32_t;
using dim_t = uint32_t;
using index_t = uint64_t;

struct Config {
  dim_t rank = 2;
  index_t max_iter = 10;
  real_t rel_tol = 1e-6;
  real_t delta = 0.5;
  index_t seed = 0;
  // "rnd" (default), "hoevd", "given" (U0 below); "hosvd" needs MKL in the
  // reference and falls back to the random factor here as it does there.
  std::string init = "rnd";
  // Extensions of this build.
  int precision = 0;            // SP_PREC_FP64 / SP_PREC_FP32 / SP_PREC_BF16
  real_t switch_tol = -1.0;     // mixed forms: |d error| that ends the
                                // reduced-precision phase (<0: automatic)
  real_t subspace_tol = 0.0;    // > 0: stop when the factor's subspace moves
                                // by less (sine of the largest principal
                                // angle) instead of on the error change
  bool verbose = true;          // per-iteration lines on stdout
  std::vector<real_t> U0;       // initial factor for init == "given"

  Config() = default;
};

class SpSymtensor;

class SymTuckerDecomp {
public:
  real_t reconstruct_error(const SpSymtensor &target) const;
  void print(std::ostream &os) const;

  // Extensions: iteration record of the run that produced this result.
  order_t order() const { return _order; }
  const std::vector<real_t> &factor() const { return _factor; }
  const std::vector<real_t> &core() const { return _core; }
  const std::vector<real_t> &errors() const { return _errors; }
  const std::vector<real_t> &core_norms() const { return _core_norms; }
  const std::vector<double> &times() const { return _times; }
  index_t iterations() const { return _errors.size(); }
  index_t fp64_sweeps() const { return _fp64_sweeps; }
  double seconds() const { return _seconds; }

  friend class SpSymtensor;

private:
  SymTuckerDecomp(const Config &config, dim_t dim, order_t order);
  void record(real_t error, real_t core_norm, std::chrono::steady_clock::time_point t0) {
    _errors.push_back(error);
    _core_norms.push_back(core_norm);
    _times.push_back(
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }

  dim_t _dim;
  dim_t _rank;
  order_t _order;

  std::vector<real_t> _core;
  std::vector<real_t> _factor;
  std::vector<real_t> _errors;
  std::vector<real_t> _core_norms;
  std::vector<double> _times;  // seconds from the loop start to each iteration's end
  index_t _fp64_sweeps = 0;
  double _seconds = 0.0;
};

class SpSymtensor {
public:
  SpSymtensor(dim_t dim, order_t order, const std::vector<dim_t> &unnz_inds,
              const std::vector<real_t> &unnzs);
  // Wraps a context created through the C-ABI (not owned).
  explicit SpSymtensor(sp_ctx *ctx);
  ~SpSymtensor();
  SpSymtensor(const SpSymtensor &) = delete;
  SpSymtensor &operator=(const SpSymtensor &) = delete;

  order_t order() const { return _order; }
  dim_t dim() const { return _dim; }
  index_t size() const { return _unnz; }

  void S3TTMc(std::vector<real_t> &res, const std::vector<real_t> &U, dim_t R,
              bool partial_sym_res = false) const;
  SymTuckerDecomp TuckerHOOI(const Config &c) const;
  SymTuckerDecomp TuckerHOQRI(const Config &c) const;
  SymTuckerDecomp TuckerAltHOOI(const Config &c) const;

  real_t norm() const { return _norm; }
  sp_ctx *context() const { return _ctx; }
  // Per-phase FunctionTimer entries of the Tucker loops (regularize, S3TTMc,
  // leftsvd, ...): each phase is timed to its end on the GPU, i.e. the host
  // waits for the stream after it (default on, as the reference's timers
  // measure each call).  Off: no timers and no waits between the phases.
  void set_phase_timing(bool on) { _phase_timing = on; }

private:
  template <class F>
  void timed(const char *name, F &&f) const;
  void init_factor(std::vector<real_t> &U, const Config &c) const;
  void hoevd_init(std::vector<real_t> &U, const Config &c) const;

  dim_t _dim = 0;
  order_t _order = 0;
  index_t _unnz = 0;
  real_t _norm = 0.0;
  sp_ctx *_ctx = nullptr;
  bool _owns = false;
  bool _phase_timing = true;
};

// Packed (lexicographic sorted multi-index) helpers.
index_t symtens_size(index_t dim, index_t order);
std::vector<real_t> symtens_coeff(size_t dim, size_t order);
void unpack_symtens(const real_t *packed, dim_t dim, order_t order, real_t *full);

// RAII wall-clock accumulator with the reference's report format.
class FunctionTimer {
public:
  explicit FunctionTimer(std::string_view name);
  ~FunctionTimer();
  static void printTimes();
  static void clear();
  void stop();

private:
  std::chrono::time_point<std::chrono::high_resolution_clock> m_start;
  std::string m_name;
  bool m_stopped = false;
};

#define START_FUNCTION_TIMER() ::symprop::FunctionTimer __ft(__FUNCTION__);

}  // namespace symprop
