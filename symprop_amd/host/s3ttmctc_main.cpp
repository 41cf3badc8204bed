// s3ttmctc: ten rounds of S3TTMc + core + factor update (one HOQRI step
// without the QR) on the MI355X path, with the interface of the reference's
// micro-driver (examples/s3ttmctc.cpp:13-93): "s3ttmctc <input_file> <R>",
// factor from std::mt19937(1) over U[1, 2), an "Insertion time" line, one
// "S3TTMcTC Time" line per round and the accumulated function times.
//
// A round is the reference's: packed Y = S3TTMc(U) into host memory,
// core = U^T Y (matrix::gemmtn), U = Y diag(W) core^T (matrix::symgemm), with
// the factor back in host memory at the end of the round.  The products run
// on the device, on the Y and U the S3TTMc left there.
#include "micro_driver.hpp"

#include <stdexcept>

namespace {

void check(int rc) {
  if (rc != SP_OK) throw std::runtime_error(sp_last_error());
}

}  // namespace

int main(int argc, char **argv) {
  using namespace symprop;
  driver::Input in;
  if (!driver::read_input(argc, argv, in)) return 1;
  std::vector<real_t> U = driver::start_factor(in.dim, in.rank);

  auto t0 = std::chrono::high_resolution_clock::now();
  SpSymtensor tensor(in.dim, in.order, in.inds, in.vals);
  std::cout << "Insertion time: " << driver::elapsed_ms(t0) << " ms" << std::endl;

  const size_t width = symtens_size(in.rank, in.order - 1);
  std::vector<real_t> core(static_cast<size_t>(in.rank) * width);
  std::vector<real_t> Y(static_cast<size_t>(in.dim) * width);
  sp_ctx *ctx = tensor.context();
  try {
    for (int round = 0; round < 10; ++round) {
      t0 = std::chrono::high_resolution_clock::now();
      tensor.S3TTMc(Y, U, in.rank, true);
      {
        FunctionTimer t("gemmtn");
        check(sp_gemmtn(ctx, core.data()));
      }
      {
        FunctionTimer t("symgemm");
        check(sp_symgemm(ctx));
        check(sp_factor_download(ctx, U.data()));
      }
      std::cout << "S3TTMcTC Time: " << driver::elapsed_ms(t0) << " ms" << std::endl;
    }
  } catch (const std::exception &e) {
    std::cerr << "s3ttmctc: " << e.what() << std::endl;
    return 1;
  }
  FunctionTimer::printTimes();
  return 0;
}
