// Thin C driver over the reference's own compressed dimension tree, built
// from the reference sources by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY (the checker's checker): it lets tests compare
// the oracle's tree arrays and S3TTMc with the reference implementation on
// the same inputs.  It uses only the reference's public interface:
//   DTree::banerjee_insert            include/detail/dtree.hpp:121-122
//   CompressedDTree(const DTree&, d)  include/detail/compressed_dtree.hpp:14
//   CompressedDTree::print            src/compressed_dtree.cpp:530-556
//   CompressedDTree::exclude_search   src/compressed_dtree.cpp:520-528
//   kronecker_product_opt5            src/compressed_dtree.cpp:421-518
// and mirrors the construction sequence of SpSymtensor's constructor
// (src/sparse/spsymtensor.cpp:16-28).

#include <detail/compressed_dtree.hpp>
#include <detail/dtree.hpp>
#include <utils/types.hpp>

#include <omp.h>

#include <cstring>
#include <iostream>
#include <sstream>
#include <span>
#include <string>
#include <vector>

using namespace symprop;

struct RefTree {
  std::unique_ptr<CompressedDTree> cdt;
  uint32_t order;
  uint32_t dim;
  std::string dump;
};

extern "C" {

RefTree *ref_build(uint32_t order, uint32_t dim, uint64_t nnz,
                   const uint32_t *inds, const double *vals) {
  auto *h = new RefTree();
  h->order = order;
  h->dim = dim;
  DTree dt;
  for (uint64_t i = 0; i < nnz; ++i) {
    dt.banerjee_insert(std::span<const dim_t>(inds + i * order, order),
                       vals[i]);
  }
  std::streambuf *old = std::cout.rdbuf();
  std::ostringstream sink;
  std::cout.rdbuf(sink.rdbuf());
  h->cdt = std::make_unique<CompressedDTree>(dt, order - 1);
  h->cdt->print();
  std::cout.rdbuf(old);
  h->dump = sink.str();
  return h;
}

// The same construction without the print() dump: what SpSymtensor's
// constructor does (src/sparse/spsymtensor.cpp:16-28), for timing it.
RefTree *ref_build_quiet(uint32_t order, uint32_t dim, uint64_t nnz,
                         const uint32_t *inds, const double *vals) {
  auto *h = new RefTree();
  h->order = order;
  h->dim = dim;
  DTree dt;
  for (uint64_t i = 0; i < nnz; ++i) {
    dt.banerjee_insert(std::span<const dim_t>(inds + i * order, order),
                       vals[i]);
  }
  h->cdt = std::make_unique<CompressedDTree>(dt, order - 1);
  return h;
}

void ref_free(RefTree *h) { delete h; }

// Text of CompressedDTree::print() for the tree built by ref_build.
const char *ref_dump(const RefTree *h) { return h->dump.c_str(); }

uint64_t ref_exclude_search(RefTree *h, const uint32_t *tuple, uint32_t len,
                            uint32_t k) {
  return h->cdt->exclude_search(std::span<const dim_t>(tuple, len), k);
}

// Packed S3TTMc: zero Y (n x ldy) and accumulate kronecker_product_opt5.
void ref_opt5(RefTree *h, const double *U, uint32_t R, double *Y,
              uint64_t ldy, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  std::vector<real_t> u(U, U + (size_t)h->dim * R);
  std::vector<real_t> y((size_t)h->dim * ldy, 0.0);
  h->cdt->kronecker_product_opt5(u, R, y, ldy);
  std::memcpy(Y, y.data(), y.size() * sizeof(double));
}

}  // extern "C"
