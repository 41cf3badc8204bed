// Allocation-free S3TTMc (sp_s3ttmc_noalloc) and the device array it uses
// as scratch.
//
// The regular entry point keeps its scratch (split-row partial sums of hub
// rows, the FP32 copy of the factor for the reduced precisions) in device
// buffers that the context grows with hipMalloc.  This form takes the same
// kernels and plan but puts that scratch in `scratch`, a file-scope device
// array of kScratchBytes that every GPU the library runs on holds until the
// process ends.  In host code the array's name only stands for the symbol;
// its device address is fetched once per device with hipGetSymbolAddress
// (noalloc_setup, called when a context is created on that device) and only
// that address is handed to kernels.
//
// Every call on a GPU shares that GPU's array and leaves its last contents
// behind.  Nothing in it has to start at zero: each partial-sum slot is
// written by its chunk before the reduction reads it, and the factor copy is
// written before the kernels read it.  Calls that overlap on one GPU (two
// contexts on different streams, or two host threads) race on the array.
#include "sp_internal.hpp"

#include <mutex>

namespace spamd {

namespace {

constexpr u64 kScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device

u64 round_up(u64 x, u64 a) { return (x + a - 1) / a * a; }

}  // namespace

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw Error(SP_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  SP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(SP_ERR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_ptr[device] = p;
}

void s3ttmc_noalloc(Ctx &c, const double *U, u32 R, double *Y, int precision) {
  if (!c.built) throw Error(SP_ERR_STATE, "tree not built");
  if (R == 0 || c.chunk_R != R || c.plan[0].h_chunks.empty())
    throw Error(SP_ERR_STATE,
                "context not prepared for this rank (run sp_s3ttmc_dev at this rank first)");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (c.device >= 0 && c.device < kMaxDevices) base = g_scratch_ptr[c.device];
  }
  if (!base) throw Error(SP_ERR_STATE, "scratch address of this device not set up");
  const u64 u32_bytes =
      precision == SP_PREC_FP64 ? 0 : round_up((u64)c.dim * R * sizeof(float), 256);
  const u64 partial_bytes = s3ttmc_partial_slots(c, precision) * c.P * sizeof(double);
  if (u32_bytes + partial_bytes > kScratchBytes)
    throw Error(SP_ERR_UNSUPPORTED,
                "scratch need of " + std::to_string(u32_bytes + partial_bytes) +
                    " bytes exceeds the 64 MiB scratch array");
  unsigned char *b = static_cast<unsigned char *>(base);
  float *U32 = precision == SP_PREC_FP64 ? nullptr : reinterpret_cast<float *>(b);
  double *partial = reinterpret_cast<double *>(b + u32_bytes);
  s3ttmc_launch(c, precision, U, U32, Y, partial, false);
}

}  // namespace spamd
