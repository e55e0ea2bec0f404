// Shared host/device definitions for the FastXC MI355X kernels.
//
// Error handling follows one rule on both sides of the C ABI: every entry
// point returns an int status (0 = success) and records a message that
// fxc_last_error() returns.  No exception crosses the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#define FXC_OK 0
#define FXC_ERR_ARG 1
#define FXC_ERR_HIP 2
#define FXC_ERR_ALLOC 3
#define FXC_ERR_UNSUPPORTED 4

namespace fxc {

void set_error(const char *fmt, ...);
const char *last_error();

// Wraps a HIP call: on failure records the HIP error string and returns
// FXC_ERR_HIP from the enclosing function.
#define FXC_HIP(call)                                                        \
  do {                                                                       \
    hipError_t fxc_e_ = (call);                                              \
    if (fxc_e_ != hipSuccess) {                                              \
      ::fxc::set_error("%s:%d %s failed: %s", __FILE__, __LINE__, #call,      \
                       hipGetErrorString(fxc_e_));                           \
      return FXC_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define FXC_CHECK_ARG(cond, ...)                                             \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::fxc::set_error(__VA_ARGS__);                                         \
      return FXC_ERR_ARG;                                                    \
    }                                                                        \
  } while (0)

// Propagates a non-zero status from a nested fxc call.
#define FXC_TRY(call)                                                        \
  do {                                                                       \
    int fxc_s_ = (call);                                                     \
    if (fxc_s_ != FXC_OK) return fxc_s_;                                     \
  } while (0)

// ---------------------------------------------------------------------------
// Workgroup FFT plan (device side).  A transform of length n = prod(radix)
// runs as in-place mixed-radix decimation-in-frequency passes in LDS.  Pass p
// splits spans of length span[p] into radix[p] sub-spans of length
// stride[p] = span[p]/radix[p]; its twiddles w^(j*q), j < stride[p],
// 1 <= q < radix[p], sit at tw[tw_off[p] + j*(radix[p]-1) + q-1] with
// w = exp(sign*2*pi*i/span[p]).  The result is left in digit-reversed order:
// bin k is read from position fft_pos(plan, k) = sum_p d_p*stride[p], where
// k = d_0 + radix[0]*(d_1 + radix[1]*(d_2 + ...)).
constexpr int kMaxPasses = 16;

struct FftPlanDev {
  int n;
  int npass;
  int radix[kMaxPasses];
  int span[kMaxPasses];
  int stride[kMaxPasses];
  int tw_off[kMaxPasses];
};

// Host-side owner of a workgroup FFT plan and its twiddle table.
struct FftPlan {
  FftPlanDev dev{};
  int sign = -1;        // -1: forward exp(-i...), +1: inverse exp(+i...)
  float2 *d_tw = nullptr;
  size_t tw_count = 0;
};

// Factor n into radices this library implements; returns false when n has
// a prime factor with no butterfly (then a Bluestein plan is needed).
bool factor_radices(int n, int *radix, int *npass);
// Builds the plan and uploads its FP64-computed twiddle table.
int fft_plan_create(FftPlan *plan, int n, int sign);
void fft_plan_destroy(FftPlan *plan);

// Device memory helpers that record an error on failure.
int dmalloc(void **ptr, size_t bytes);
void dfree(void *ptr);

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace fxc
