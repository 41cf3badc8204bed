// Shared pieces of the S3TTMc micro-drivers (symtens_ttmc, s3ttmctc): the
// command line "<input_file> <R>", the tensor read with the reference's .tns
// semantics, and the factor the reference's drivers start from.
#pragma once

#include "symtucker.hpp"

#include "../../include/symprop_amd.h"

#include <chrono>
#include <cstdint>
#include <iostream>
#include <random>
#include <string>
#include <vector>

namespace symprop::driver {

struct Input {
  std::string path;
  dim_t rank = 0;
  order_t order = 0;
  dim_t dim = 0;
  std::vector<dim_t> inds;
  std::vector<real_t> vals;
};

// Parses "<input_file> <R>" and reads the tensor (header "order dim", then
// sorted tuples until the first failed read).  Prints the reference's
// messages and returns false on a usage or open error.
inline bool read_input(int argc, char **argv, Input &in) {
  if (argc < 3) {
    std::cerr << "Usage: " << argv[0] << " <input_file> <R>" << std::endl;
    return false;
  }
  in.path = argv[1];
  in.rank = static_cast<dim_t>(std::stoi(argv[2]));
  uint64_t nnz = 0;
  if (sp_load_tns(in.path.c_str(), &in.order, &in.dim, &nnz, nullptr, nullptr) != SP_OK) {
    std::cerr << "Error opening file: " << in.path << std::endl;
    return false;
  }
  in.inds.resize(static_cast<size_t>(nnz) * in.order);
  in.vals.resize(nnz);
  if (sp_load_tns(in.path.c_str(), &in.order, &in.dim, &nnz, in.inds.data(),
                  in.vals.data()) != SP_OK) {
    std::cerr << "Error reading file: " << in.path << ": " << sp_last_error() << std::endl;
    return false;
  }
  return true;
}

// dim x R factor with entries drawn from U[1, 2) by std::mt19937 seeded with
// 1, in row-major order (the drivers' "random initialization").
inline std::vector<real_t> start_factor(size_t dim, size_t R) {
  std::vector<real_t> U(dim * R);
  std::mt19937 gen(1);
  std::uniform_real_distribution<real_t> draw(1.0, 2.0);
  for (auto &u : U) u = draw(gen);
  return U;
}

inline long long elapsed_ms(std::chrono::high_resolution_clock::time_point t0) {
  return std::chrono::duration_cast<std::chrono::milliseconds>(
             std::chrono::high_resolution_clock::now() - t0)
      .count();
}

}  // namespace symprop::driver
