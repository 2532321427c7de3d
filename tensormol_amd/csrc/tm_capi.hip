// Library-wide entry points: version and the thread-local last-error string.
#include <stdarg.h>
#include <stdio.h>

#include "tm_common.h"

namespace {
thread_local char g_err[512] = {0};
}

void tm_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" {

int tm_version(void) { return 1; }

const char* tm_last_error(void) { return g_err; }

// Clear an int32 device array with a kernel (usable inside captured graphs,
// where memset nodes are avoided).
__global__ void tm_zero_i32_kernel(int n, int* p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

int tm_zero_i32(int n, int* p, void* stream) {
  if (n <= 0) return TM_OK;
  tm_zero_i32_kernel<<<tm_div_up(n, 256), 256, 0, (hipStream_t)stream>>>(n, p);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
