// Host-side core: error strings, device allocation, FFT plan construction.
#include "fxc_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <vector>

namespace fxc {

static thread_local char g_err[1024] = {0};

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *last_error() { return g_err; }

int dmalloc(void **ptr, size_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) return FXC_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    return FXC_ERR_ALLOC;
  }
  return FXC_OK;
}

void dfree(void *ptr) {
  if (ptr) (void)hipFree(ptr);
}

// Radix preference: large register-resident butterflies first keeps the LDS
// pass count low (18522 = 21*21*21*2, 18000 = 16*15*15*5, 18144 = 21*16*9*6);
// the odd primes 23 and 29 cover the correlation length 2001 = 3*23*29.
bool factor_radices(int n, int *radix, int *npass, int max_radix, bool largest_last) {
  static const int prefs[] = {21, 16, 15, 14, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2, 23, 29};
  int m = n;
  int cnt = 0;
  if (n < 1) return false;
  for (int r : prefs) {
    if (r > max_radix && r != 23 && r != 29) continue;
    while (m % r == 0 && m > 1) {
      if (cnt >= kMaxPasses) return false;
      radix[cnt++] = r;
      m /= r;
    }
  }
  if (m != 1) return false;
  if (cnt == 0) {  // n == 1: a single trivial pass is not needed
    *npass = 0;
    return true;
  }
  if (largest_last) std::sort(radix, radix + cnt);
  *npass = cnt;
  return true;
}

int fft_plan_create(FftPlan *plan, int n, int sign, int max_radix, bool largest_last) {
  FftPlanDev d{};
  d.n = n;
  if (!factor_radices(n, d.radix, &d.npass, max_radix, largest_last)) {
    set_error("fft length %d has a prime factor without a butterfly", n);
    return FXC_ERR_UNSUPPORTED;
  }
  auto magic = [](int d) -> unsigned {
    return d <= 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned long long)d - 1) / (unsigned long long)d);
  };
  if (n >= (1 << 16)) {
    set_error("fft length %d exceeds the workgroup FFT limit", n);
    return FXC_ERR_UNSUPPORTED;
  }
  int span = n;
  for (int p = 0; p < d.npass; ++p) {
    const int r = d.radix[p];
    const int stride = span / r;
    d.span[p] = span;
    d.stride[p] = stride;
    d.tstep[p] = n / span;
    d.smagic[p] = magic(stride);
    d.bmagic[p] = magic(n / r);
    d.rmagic[p] = magic(r);
    span = stride;
  }
  d.nmagic = magic(n);
  d.ncoarse = (n + kTwFine - 1) / kTwFine;
  // two-level table: fine[u] = w^u (u < 128), coarse[v] = w^(128 v)
  std::vector<float2> tw(kTwFine + d.ncoarse);
  for (int u = 0; u < kTwFine; ++u) {
    const double a = sign * 2.0 * M_PI * (double)(u % n) / (double)n;
    tw[u] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int v = 0; v < d.ncoarse; ++v) {
    const long long e = ((long long)v * kTwFine) % n;
    const double a = sign * 2.0 * M_PI * (double)e / (double)n;
    tw[kTwFine + v] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  plan->dev = d;
  plan->sign = sign;
  plan->tw_count = tw.size();
  plan->d_tw = nullptr;
  if (!tw.empty()) {
    FXC_TRY(dmalloc((void **)&plan->d_tw, tw.size() * sizeof(float2)));
    FXC_HIP(hipMemcpy(plan->d_tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
  }
  return FXC_OK;
}

void fft_plan_destroy(FftPlan *plan) {
  if (!plan) return;
  dfree(plan->d_tw);
  plan->d_tw = nullptr;
}

}  // namespace fxc

extern "C" const char *fxc_last_error(void) { return fxc::last_error(); }
