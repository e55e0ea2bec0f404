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

#include "fastxc_amd.h"  // status codes FXC_OK, FXC_ERR_*

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
// stride[p] = span[p]/radix[p]; butterfly j of a span is twiddled by
// w_n^(j*q*tstep[p]) (tstep = n/span, w_n = exp(sign*2*pi*i/n)).  The result
// is left in digit-reversed order: bin k is at fft_pos(plan, k) =
// sum_p d_p*stride[p], where k = d_0 + radix[0]*(d_1 + radix[1]*(d_2 + ...)).
// Twiddles come from a two-level table w_n^e = coarse[e >> 7] * fine[e & 127]
// (fine: 128 entries, coarse: ceil(n/128)), computed in FP64 on the host and
// staged in LDS.  Integer divisions use multiply-high by a precomputed
// ceil(2^32/d), exact for operands below 2^16.
constexpr int kMaxPasses = 16;
constexpr int kTwFine = 128;

struct FftPlanDev {
  int n;
  int npass;
  int ncoarse;
  unsigned nmagic;             // ceil(2^32/n), 0 when n == 1
  int radix[kMaxPasses];
  int span[kMaxPasses];
  int stride[kMaxPasses];
  int tstep[kMaxPasses];
  unsigned smagic[kMaxPasses];  // ceil(2^32/stride), 0 when stride == 1
  unsigned bmagic[kMaxPasses];  // ceil(2^32/(n/radix))
  unsigned rmagic[kMaxPasses];  // ceil(2^32/radix)
};

// float2 entries of the LDS twiddle region of a length-n plan.
inline __host__ __device__ int tw_lds_count(int n) { return kTwFine + (n + kTwFine - 1) / kTwFine; }

// Host-side owner of a workgroup FFT plan and its twiddle table.
struct FftPlan {
  FftPlanDev dev{};
  int sign = -1;        // -1: forward exp(-i...), +1: inverse exp(+i...)
  float2 *d_tw = nullptr;
  size_t tw_count = 0;
};

// Factor n into radices this library implements; returns false when n has
// a prime factor with no butterfly (then a Bluestein plan is needed).
// Radices above max_radix are not used (kernels with tight register budgets
// ask for small butterflies).
// largest_last: the radices in ascending order, so the last pass (stride 1)
// has the largest butterfly (used when a consumer folds the last pass into
// its output step).
bool factor_radices(int n, int *radix, int *npass, int max_radix = 32, bool largest_last = false);
// Builds the plan and uploads its FP64-computed twiddle table.
int fft_plan_create(FftPlan *plan, int n, int sign, int max_radix = 32, bool largest_last = false);
void fft_plan_destroy(FftPlan *plan);

// Device memory helpers that record an error on failure.
int dmalloc(void **ptr, size_t bytes);
void dfree(void *ptr);

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace fxc
