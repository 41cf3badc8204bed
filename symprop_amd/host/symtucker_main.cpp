// symtucker command-line driver on the MI355X S3TTMc path.
//
// Interface kept from the reference driver (examples/symtucker.cpp:29-139):
// the flags -f -i -r -t -d -a -s -o -init -h, the usage line, the
// "Option X is not defined" message, exit status 1 for -h and for unknown
// flags, the "Insertion time" and per-algorithm "Running"/"Time" lines, the
// output sinks ("" -> discarded, "stdout", "stderr", else a file path), and
// -a -1 (default) running HOOI, HOQRI and AltHOOI in turn.  Extension flags:
// -p <0|1|2> selects FP64 / FP32 / BF16 S3TTMc iterations (the reduced forms
// finish with FP64 iterations); -c <file> reads the tensor through a binary
// tuple cache (written on first use, reused while the .tns file is unchanged).
//
// Design: options are rows of a table (flag, value parser), algorithms are
// rows of a second table (id, printed name, member function), so adding a
// flag or an algorithm is one line.
#include "symtucker.hpp"

#include "../../include/symprop_amd.h"

#include <array>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace {

using symprop::Config;
using symprop::FunctionTimer;
using symprop::SpSymtensor;
using symprop::SymTuckerDecomp;

struct Invocation {
  Config config;
  std::string tensor_path;
  std::string cache_path;  // binary tuple cache (optional)
  std::string sink;        // where the decomposition is written
  int alg = -1;      // -1: all algorithms
};

using Setter = void (*)(Invocation &, const char *);

struct OptionRow {
  std::string_view flag;
  Setter set;
};

// Numeric conversions follow the reference: -i goes through a double, the
// integer flags through int.
const std::array<OptionRow, 11> kOptions{{
    {"-f", [](Invocation &v, const char *s) { v.tensor_path = s; }},
    {"-i", [](Invocation &v, const char *s) {
       v.config.max_iter = static_cast<symprop::index_t>(std::stod(s));
     }},
    {"-r", [](Invocation &v, const char *s) {
       v.config.rank = static_cast<symprop::dim_t>(std::stoi(s));
     }},
    {"-t", [](Invocation &v, const char *s) { v.config.rel_tol = std::stod(s); }},
    {"-d", [](Invocation &v, const char *s) { v.config.delta = std::stod(s); }},
    {"-a", [](Invocation &v, const char *s) { v.alg = std::stoi(s); }},
    {"-s", [](Invocation &v, const char *s) {
       v.config.seed = static_cast<symprop::index_t>(std::stoi(s));
     }},
    {"-o", [](Invocation &v, const char *s) { v.sink = s; }},
    {"-init", [](Invocation &v, const char *s) { v.config.init = s; }},
    {"-p", [](Invocation &v, const char *s) { v.config.precision = std::stoi(s); }},
    {"-c", [](Invocation &v, const char *s) { v.cache_path = s; }},
}};

struct AlgorithmRow {
  int id;
  const char *name;
  SymTuckerDecomp (SpSymtensor::*run)(const Config &) const;
};

const std::array<AlgorithmRow, 3> kAlgorithms{{
    {0, "HOOI", &SpSymtensor::TuckerHOOI},
    {1, "HOQRI", &SpSymtensor::TuckerHOQRI},
    {2, "AltHOOI", &SpSymtensor::TuckerAltHOOI},
}};

void print_usage(const char *prog) {
  std::cout << prog << " -f <input_file> -i <max_iter> -r <rank> -t <rel_tol> -d <delta>"
            << " -a <alg> -s <seed> -o <output_file> -init <init method>" << std::endl;
}

// Returns false when the driver must stop with status 1 (-h or a flag that
// is unknown or lacks its value).
bool read_arguments(int argc, char **argv, Invocation &inv) {
  int i = 1;
  while (i < argc) {
    const std::string_view word = argv[i];
    if (word == "-h") {
      print_usage(argv[0]);
      return false;
    }
    const OptionRow *row = nullptr;
    for (const auto &r : kOptions)
      if (r.flag == word) row = &r;
    if (row == nullptr || i + 1 >= argc) {
      std::cout << "Option " << argv[i] << " is not defined\n";
      print_usage(argv[0]);
      return false;
    }
    row->set(inv, argv[i + 1]);
    i += 2;
  }
  return true;
}

std::string sink_path(const std::string &sink) {
  if (sink.empty()) return "/dev/null";
  if (sink == "stdout" || sink == "stderr") return "/dev/" + sink;
  return sink;
}

long long elapsed_ms(std::chrono::high_resolution_clock::time_point since) {
  return std::chrono::duration_cast<std::chrono::milliseconds>(
             std::chrono::high_resolution_clock::now() - since)
      .count();
}

// Timed run of one algorithm, then the accumulated function times.
SymTuckerDecomp run_one(const SpSymtensor &t, const AlgorithmRow &a, const Config &c) {
  std::cout << "Running " << a.name << " algorithm" << std::endl;
  const auto t0 = std::chrono::high_resolution_clock::now();
  SymTuckerDecomp result = (t.*a.run)(c);
  std::cout << "---------- Time: " << elapsed_ms(t0) << " ms ----------" << std::endl;
  FunctionTimer::printTimes();
  FunctionTimer::clear();
  return result;
}

// The tensor file through the library's loader (two calls: shape, then data).
void load_tensor(const std::string &path, const std::string &cache, symprop::order_t &order,
                 symprop::dim_t &dim, std::vector<symprop::dim_t> &inds,
                 std::vector<symprop::real_t> &vals) {
  const char *cp = cache.empty() ? nullptr : cache.c_str();
  auto call = [&](uint64_t &nnz, uint32_t *i, double *v) {
    return cp ? sp_load_tns_cached(path.c_str(), cp, &order, &dim, &nnz, i, v)
              : sp_load_tns(path.c_str(), &order, &dim, &nnz, i, v);
  };
  uint64_t nnz = 0;
  if (call(nnz, nullptr, nullptr) != SP_OK) {
    std::cerr << "Error opening file: " << path << std::endl;
    throw std::runtime_error("Error opening file in SymtensLoader");
  }
  inds.resize(static_cast<size_t>(nnz) * order);
  vals.resize(nnz);
  if (call(nnz, inds.data(), vals.data()) != SP_OK)
    throw std::runtime_error(std::string("symtucker: ") + sp_last_error());
}

}  // namespace

int main(int argc, char **argv) {
  Invocation inv;
  if (!read_arguments(argc, argv, inv)) return 1;

  symprop::order_t order = 0;
  symprop::dim_t dim = 0;
  std::vector<symprop::dim_t> inds;
  std::vector<symprop::real_t> vals;
  load_tensor(inv.tensor_path, inv.cache_path, order, dim, inds, vals);

  const auto t0 = std::chrono::high_resolution_clock::now();
  const SpSymtensor tensor(dim, order, inds, vals);
  std::cout << "Insertion time: " << elapsed_ms(t0) << " ms" << std::endl;
  FunctionTimer::printTimes();
  FunctionTimer::clear();

  std::ofstream out(sink_path(inv.sink));
  for (const auto &a : kAlgorithms)
    if (inv.alg == -1 || inv.alg == a.id) run_one(tensor, a, inv.config).print(out);
  return 0;
}
