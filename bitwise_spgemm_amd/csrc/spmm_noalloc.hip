// spmm_noalloc.hip — the FP32 0/1 product for callers that cannot allocate
// (code under graph capture, tight serving loops, harnesses that forbid
// allocation): bws_spmm_fp32_noalloc is bws_spmm at FP32 with the mask stream
// of the streamed kernel kept in a file-scope device array instead of a
// caller workspace.
//
// The array, `scratch`, is kScratchBytes (16 MiB) on every device the library
// runs on and lives until the process ends.  In host code the array's name is
// a stand-in, not a device address, so it is never passed to a kernel: on the
// first call on a device, with that device current, hipGetSymbolAddress gives
// that device's address of its own copy, which is kept per device and used
// from then on.  Every call writes the whole stream it reads (nothing has to
// start at zero); calls that overlap on one device (several streams or host
// threads) race on the array.
#include <mutex>

#include "bws_internal.hpp"

namespace bws {
namespace {

constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
bws_status scratch_for_current_device(void** out) {
  int dev = -1;
  BWS_TRY(hip_check(hipGetDevice(&dev)));
  if (dev < 0 || dev >= kMaxDevices) return BWS_ENODEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    BWS_TRY(hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))));
    if (!p) return BWS_EHIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BWS_OK;
}

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

size_t bws_spmm_fp32_noalloc_max_bytes(void) { return kScratchBytes; }

bws_status bws_spmm_fp32_noalloc(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N,
                                 float* C, int64_t ldc, float alpha, float beta,
                                 bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (A->values) return BWS_EUNSUPPORTED;
  if (N < 0) return BWS_EINVAL;
  if (mask_stream_bytes(A->rows, A->cols) > kScratchBytes) return BWS_ECAPACITY;
  void* ws = nullptr;
  BWS_TRY(scratch_for_current_device(&ws));
  return bws_spmm(A, B, BWS_DTYPE_F32, ldb, N, C, ldc, alpha, beta, BWS_PREC_FP32, nullptr, ws,
                  kScratchBytes, stream);
}

}  // extern "C"
