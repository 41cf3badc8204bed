// symtens_ttmc: one packed S3TTMc of a .tns tensor on the MI355X path, with
// the interface of the reference's micro-driver (examples/symtens_ttmc.cpp:
// 11-88): "symtens_ttmc <input_file> <R>", factor from std::mt19937(1) over
// U[1, 2), "Insertion time" and "S3TTMc Time" lines, and the packed Y
// (dim rows of C(R+N-2, N-1) values, fixed with 5 decimals, each followed by
// ", ") written to ttmc.out in the working directory.
#include "micro_driver.hpp"

#include <cstdio>
#include <memory>

int main(int argc, char **argv) {
  using namespace symprop;
  driver::Input in;
  if (!driver::read_input(argc, argv, in)) return 1;
  const std::vector<real_t> U = driver::start_factor(in.dim, in.rank);

  auto t0 = std::chrono::high_resolution_clock::now();
  SpSymtensor tensor(in.dim, in.order, in.inds, in.vals);
  std::cout << "Insertion time: " << driver::elapsed_ms(t0) << " ms" << std::endl;

  t0 = std::chrono::high_resolution_clock::now();
  const size_t width = symtens_size(in.rank, in.order - 1);
  std::vector<real_t> Y(static_cast<size_t>(in.dim) * width);
  tensor.S3TTMc(Y, U, in.rank, true);
  std::cout << "S3TTMc Time: " << driver::elapsed_ms(t0) << " ms" << std::endl;

  std::unique_ptr<std::FILE, int (*)(std::FILE *)> out(std::fopen("ttmc.out", "w"), std::fclose);
  if (!out) {
    std::cerr << "Error opening ttmc.out" << std::endl;
    return 1;
  }
  for (size_t i = 0; i < in.dim; ++i) {
    const real_t *row = Y.data() + i * width;
    for (size_t j = 0; j < width; ++j) std::fprintf(out.get(), "%.5f, ", row[j]);
    std::fputc('\n', out.get());
  }
  return 0;
}
