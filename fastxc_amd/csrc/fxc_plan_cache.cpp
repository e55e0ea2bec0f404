#include "fxc_plan_cache.hpp"

#include "fxc_fft_device.hpp"
#include "fxc_fft_large.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

namespace fxc {

namespace {
std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, std::unique_ptr<FftPlan>> g_cplx;
std::map<std::tuple<int, int, int, int, bool>, std::unique_ptr<RealFftPlan>> g_real;
std::map<std::tuple<int, int>, std::unique_ptr<BluesteinPlan>> g_blue;

constexpr int kBlueThreads = 256;

// B_s = DFT_m(b_s) / m, b_s[d] = conj(c_s(d)) for |d| < n (indices mod m),
// 0 elsewhere; block 0 builds s = -1, block 1 s = +1.  Output keeps the
// forward transform's digit-reversed layout.
__global__ void __launch_bounds__(kBlueThreads)
bluestein_bhat_kernel(const float2 *__restrict__ chirp, int n, FftPlanDev plan,
                      const float2 *__restrict__ tw, float2 *__restrict__ out_f,
                      float2 *__restrict__ out_i) {
  extern __shared__ float2 lds[];
  const int m = plan.n;
  float2 *twl = lds + m;
  for (int i = threadIdx.x; i < kTwFine + plan.ncoarse; i += blockDim.x) twl[i] = tw[i];
  const bool fwd = blockIdx.x == 0;
  for (int d = threadIdx.x; d < m; d += blockDim.x) {
    float2 v = make_float2(0.f, 0.f);
    if (d < n)
      v = chirp[d];
    else if (d > m - n)
      v = chirp[m - d];
    // b_{-1} = conj(c_{-1}) ; b_{+1} = conj(c_{+1}) = c_{-1}
    if (fwd) v.y = -v.y;
    lds[d] = v;
  }
  __syncthreads();
  lds_fft<kRadixAll>(lds, 1, plan, twl, -1.f);
  float2 *o = fwd ? out_f : out_i;
  const float inv_m = 1.f / (float)m;
  for (int i = threadIdx.x; i < m; i += blockDim.x) o[i] = cscale(lds[i], inv_m);
}

void bluestein_destroy(BluesteinPlan *p) {
  dfree(p->d_chirp);
  dfree(p->d_bhat_f);
  dfree(p->d_bhat_i);
  p->d_chirp = p->d_bhat_f = p->d_bhat_i = nullptr;
}

int bluestein_create(BluesteinPlan *p, int n) {
  p->n = n;
  p->m = bluestein_length(n);
  FXC_CHECK_ARG(p->m > 0, "bluestein: no convolution length for n=%d", n);
  const int m = p->m;
  FXC_CHECK_ARG(m + tw_lds_count(m) <= kLdsComplexCapacity, "bluestein: n=%d too long", n);
  std::vector<float2> chirp((size_t)n);
  for (int j = 0; j < n; ++j) {
    const long long e = ((long long)j * j) % (2LL * n);  // phase pi*e/n, exact reduction
    const double a = M_PI * (double)e / (double)n;
    chirp[j] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  FXC_TRY(dmalloc((void **)&p->d_chirp, (size_t)n * sizeof(float2)));
  FXC_TRY(dmalloc((void **)&p->d_bhat_f, (size_t)m * sizeof(float2)));
  FXC_TRY(dmalloc((void **)&p->d_bhat_i, (size_t)m * sizeof(float2)));
  FXC_HIP(hipMemcpy(p->d_chirp, chirp.data(), (size_t)n * sizeof(float2), hipMemcpyHostToDevice));
  const size_t lds = ((size_t)m + tw_lds_count(m)) * sizeof(float2);
  FXC_HIP(hipFuncSetAttribute((const void *)bluestein_bhat_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(bluestein_bhat_kernel, dim3(2), dim3(kBlueThreads), lds, 0, p->d_chirp, n,
                     p->fwd->dev, p->fwd->d_tw, p->d_bhat_f, p->d_bhat_i);
  FXC_HIP(hipGetLastError());
  FXC_HIP(hipDeviceSynchronize());
  return FXC_OK;
}
}  // namespace

int bluestein_length(int n) {
  int radix[kMaxPasses], np = 0;
  for (int m = std::max(2 * n - 1, 1); m <= 65536; ++m)
    if (factor_radices(m, radix, &np)) return m;
  return 0;
}

int plan_cache_bluestein(int n, const BluesteinPlan **out, bool create) {
  const int m = bluestein_length(n);
  FXC_CHECK_ARG(m > 0, "bluestein: no convolution length for n=%d", n);
  const FftPlan *fwd = nullptr, *inv = nullptr;
  FXC_TRY(plan_cache_complex(m, -1, &fwd, 32, create));
  FXC_TRY(plan_cache_complex(m, 1, &inv, 32, create));
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, n);
  auto it = g_blue.find(key);
  if (it == g_blue.end()) {
    FXC_CHECK_ARG(create, "bluestein plan of length %d is not prepared on device %d", n, dev);
    std::unique_ptr<BluesteinPlan> p(new BluesteinPlan());
    p->fwd = fwd;
    p->inv = inv;
    const int rc = bluestein_create(p.get(), n);
    if (rc != FXC_OK) {
      bluestein_destroy(p.get());
      return rc;
    }
    it = g_blue.emplace(key, std::move(p)).first;
  }
  *out = it->second.get();
  return FXC_OK;
}

int plan_cache_complex(int n, int sign, const FftPlan **out, int max_radix, bool create) {
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, n, sign, max_radix);
  auto it = g_cplx.find(key);
  if (it == g_cplx.end()) {
    FXC_CHECK_ARG(create, "FFT plan of length %d is not prepared on device %d", n, dev);
    std::unique_ptr<FftPlan> p(new FftPlan());
    FXC_TRY(fft_plan_create(p.get(), n, sign, max_radix));
    it = g_cplx.emplace(key, std::move(p)).first;
  }
  *out = it->second.get();
  return FXC_OK;
}

int plan_cache_real(int n, int sign, const RealFftPlan **out, int max_radix, bool largest_last) {
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, n, sign, max_radix, largest_last);
  auto it = g_real.find(key);
  if (it == g_real.end()) {
    std::unique_ptr<RealFftPlan> p(new RealFftPlan());
    FXC_TRY(real_fft_plan_create(p.get(), n, sign, max_radix, largest_last));
    it = g_real.emplace(key, std::move(p)).first;
  }
  *out = it->second.get();
  return FXC_OK;
}

void plan_cache_release() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_blue) bluestein_destroy(kv.second.get());
  g_blue.clear();
  for (auto &kv : g_cplx) fft_plan_destroy(kv.second.get());
  for (auto &kv : g_real) real_fft_plan_destroy(kv.second.get());
  g_cplx.clear();
  g_real.clear();
}

}  // namespace fxc

extern "C" void fxc_release_plans(void) {
  fxc::large_release();
  fxc::plan_cache_release();
}
