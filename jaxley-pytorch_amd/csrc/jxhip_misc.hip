// Host-only helpers of the C ABI (no GPU call).
#include <hip/hip_runtime.h>

#include "../../include/jxhip.h"

extern "C" {

int jxhip_version(void) { return 100; /* 0.1.0 */ }

const char* jxhip_error_string(int code) {
  return hipGetErrorString(static_cast<hipError_t>(code));
}

}  // extern "C"
