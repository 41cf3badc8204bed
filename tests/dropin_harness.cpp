// Calls the host C++ drop-in (symprop_amd/host/symtucker.hpp) the way the
// reference's callers do: SpSymtensor from tuple/value vectors
// (include/detail/sparse/spsymtensor.hpp:26-65), then S3TTMc with the packed
// result (partial_sym_res = true) and with unpacked rows (the default,
// spsymtensor.cpp:30-110, as HOOI calls it at spsymtensor.cpp:223), and a
// few HOOI iterations from a given factor.
//
// usage: dropin_harness <input> <output>
//   input  (binary): u32 dim, order, R, iters; u64 nnz; u32 inds[nnz*order];
//                    f64 vals[nnz]; f64 U[dim*R]
//   output (binary): f64 packed[dim*P]; f64 full[dim*R^(order-1)];
//                    f64 errors[iters]; f64 factor[dim*R]
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "symtucker.hpp"

namespace {

template <class T>
void get(std::FILE *f, T *p, size_t n) {
  if (std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input");
}

template <class T>
void put(std::FILE *f, const std::vector<T> &v) {
  if (std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())
    throw std::runtime_error("short write");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
    return 2;
  }
  try {
    std::FILE *in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input");
    uint32_t hdr[4];
    uint64_t nnz = 0;
    get(in, hdr, 4);
    get(in, &nnz, 1);
    const uint32_t dim = hdr[0], order = hdr[1], R = hdr[2], iters = hdr[3];
    std::vector<symprop::dim_t> inds(nnz * order);
    std::vector<double> vals(nnz), U((size_t)dim * R);
    get(in, inds.data(), inds.size());
    get(in, vals.data(), vals.size());
    get(in, U.data(), U.size());
    std::fclose(in);

    symprop::SpSymtensor t(dim, order, inds, vals);
    std::vector<double> packed, full;
    t.S3TTMc(packed, U, R, true);
    t.S3TTMc(full, U, R);
    symprop::Config c;
    c.rank = R;
    c.max_iter = iters;
    c.rel_tol = 0.0;
    c.delta = 0.5;
    c.init = "given";
    c.U0 = U;
    c.verbose = false;
    const symprop::SymTuckerDecomp d = t.TuckerHOOI(c);

    std::FILE *out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    put(out, packed);
    put(out, full);
    put(out, d.errors());
    put(out, d.factor());
    std::fclose(out);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "dropin_harness: %s\n", e.what());
    return 1;
  }
  return 0;
}
