// status.cpp — status strings, last HIP error, ABI version, shared validation.
#include "bws_internal.hpp"

namespace bws {

static thread_local int32_t g_last_hip_error = 0;

void set_last_hip_error(int32_t e) { g_last_hip_error = e; }

bws_status check_bitmap(const bws_bitmap* A, bool need_row_ptr) {
  if (!A) return BWS_EINVAL;
  if (!dims_ok(A->rows, A->cols)) return BWS_EINVAL;
  if (!ldw_ok(A->cols, A->ldw)) return BWS_EINVAL;
  if (A->rows > 0 && A->cols > 0 && !A->words) return BWS_EINVAL;
  if (A->values && !A->row_ptr) return BWS_EINVAL;
  if (need_row_ptr && !A->row_ptr) return BWS_EINVAL;
  if ((reinterpret_cast<uintptr_t>(A->words) & 3u) != 0) return BWS_EINVAL;
  return BWS_OK;
}

}  // namespace bws

extern "C" {

const char* bws_status_string(bws_status status) {
  switch (status) {
    case BWS_OK: return "BWS_OK";
    case BWS_EINVAL: return "BWS_EINVAL: invalid argument";
    case BWS_ECAPACITY: return "BWS_ECAPACITY: size exceeds a documented capacity";
    case BWS_EUNSUPPORTED: return "BWS_EUNSUPPORTED: unsupported dtype/precision combination";
    case BWS_EHIP: return "BWS_EHIP: a HIP call failed";
    case BWS_EWORKSPACE: return "BWS_EWORKSPACE: workspace missing, too small or not 256-byte aligned";
    case BWS_ENODEVICE: return "BWS_ENODEVICE: no usable gfx950 device";
    case BWS_ENOTFOUND: return "BWS_ENOTFOUND: no cached autotuner entry";
    default: return "BWS_UNKNOWN";
  }
}

int32_t bws_last_hip_error(void) { return bws::g_last_hip_error; }

int32_t bws_version(void) { return (1 << 16) | 0; }

}  // extern "C"
