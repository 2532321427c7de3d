// Shared helpers for the TensorMol-AMD HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TM_WAVE 64

// status codes returned by every extern "C" entry point
enum {
  TM_OK = 0,
  TM_ERR_ARG = -1,
  TM_ERR_HIP = -2,
  TM_ERR_CAPACITY = -3,
};

// flag word layout written by kernels (int32 device array, caller-zeroed)
enum {
  TM_FLAG_NB_OVERFLOW = 0,   // a row of the 4.6 A list exceeded its stride
  TM_FLAG_ANG_OVERFLOW = 1,  // a row of the 3.1 A sub-list exceeded its stride
  TM_FLAG_CELL_OVERFLOW = 2, // an atom fell outside the open-boundary cell grid (clamped)
  TM_FLAG_NONFINITE = 3,     // a non-finite coordinate was seen
  TM_NFLAGS = 8,
};

void tm_set_error(const char* fmt, ...);

#define TM_HIP_CHECK(expr)                                                     \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      tm_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,                  \
                   hipGetErrorString(_e));                                     \
      return TM_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

#define TM_LAUNCH_CHECK()                                                      \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) {                                                    \
      tm_set_error("%s:%d launch: %s", __FILE__, __LINE__,                     \
                   hipGetErrorString(_e));                                     \
      return TM_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

#define TM_REQUIRE(cond, ...)                                                  \
  do {                                                                         \
    if (!(cond)) {                                                             \
      tm_set_error(__VA_ARGS__);                                               \
      return TM_ERR_ARG;                                                       \
    }                                                                          \
  } while (0)

static inline unsigned tm_div_up(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Maximum number of elements a model may carry (descriptor offsets are sized by it).
#define TM_MAX_ELE 8

// Minimum-image displacement along one axis for an orthorhombic box edge L
// (L <= 0 means open boundary).  Evaluated in float64 with the product and
// the difference rounded separately (no fused multiply-add), as the host
// restatement of the neighbour predicate does.
__device__ __forceinline__ double tm_min_image(double d, double L) {
#pragma clang fp contract(off)
  if (L > 0.0) d = d - L * rint(d / L);
  return d;
}

// wave-level sum of a float (butterfly over 64 lanes)
__device__ __forceinline__ float tm_wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double tm_wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
