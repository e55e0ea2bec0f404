// Batched FFT entry points (complex and real) over the workgroup LDS FFT.
//
// These replace the cuFFT plans of the reference stages: R2C/C2R of the
// window, prefilter and output lengths in sac2spec
// (native/sac2spec/device_memory.cu:23-32), the C2R of xc_fast
// (native/xc/memory/memory.cu:478-483) and the C2C plans of ncf_pws /
// ncf_tfpws (native/pws/worker/gpu.cu:197-227,
// native/tfpws/core/tfpws_workspace.cu:190-235).  Conventions follow the
// DFT definition: forward exp(-i...), inverse exp(+i...), both unnormalised,
// R2C returns bins 0..n/2.
#include "fxc_fft_large.hpp"
#include "fxc_fft_real.hpp"
#include "fxc_plan_cache.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fxc {

int real_fft_plan_create(RealFftPlan *plan, int n, int sign, int max_radix, bool largest_last) {
  if (n < 2 || (n & 1)) {
    set_error("real fft length %d must be even and >= 2", n);
    return FXC_ERR_ARG;
  }
  plan->n = n;
  FXC_TRY(fft_plan_create(&plan->half, n / 2, sign, max_radix, largest_last));
  const int M = n / 2;
  std::vector<float2> fold(fold_lds_count(M));
  for (int u = 0; u < kTwFine; ++u) {
    const double a = sign * 2.0 * M_PI * (double)u / (double)n;
    fold[u] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int v = 0; v + kTwFine < (int)fold.size(); ++v) {
    const double a = sign * 2.0 * M_PI * (double)((long long)v * kTwFine % n) / (double)n;
    fold[kTwFine + v] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  plan->fold_count = (int)fold.size();
  FXC_TRY(dmalloc((void **)&plan->d_fold, fold.size() * sizeof(float2)));
  FXC_HIP(hipMemcpy(plan->d_fold, fold.data(), fold.size() * sizeof(float2), hipMemcpyHostToDevice));
  return FXC_OK;
}

void real_fft_plan_destroy(RealFftPlan *plan) {
  fft_plan_destroy(&plan->half);
  dfree(plan->d_fold);
  plan->d_fold = nullptr;
}

int max_transforms_per_wg(const FftPlanDev &plan, int nthreads, int lds_complex_capacity) {
  (void)nthreads;  // threads loop over butterflies; only LDS capacity limits
  const int room = lds_complex_capacity - tw_lds_count(plan.n);
  return room / (plan.n > 0 ? plan.n : 1);
}

constexpr int kFftThreads = 1024;
constexpr int kFftThreadsOddPrime = 256;

template <unsigned kMask, int kThreads>
__global__ void __launch_bounds__(kThreads)
c2c_batch_kernel(const float2 *__restrict__ in, long long in_dist, float2 *__restrict__ out,
                 long long out_dist, long long ntr, int per_wg, FftPlanDev plan,
                 const float2 *__restrict__ tw, float sgn, float scale) {
  extern __shared__ float2 lds[];
  const long long t0 = (long long)blockIdx.x * per_wg;
  const int nt = (int)((ntr - t0) < per_wg ? (ntr - t0) : per_wg);
  const int n = plan.n;
  float2 *twl = lds + (size_t)per_wg * n;
  for (int i = threadIdx.x; i < kTwFine + plan.ncoarse; i += blockDim.x) twl[i] = tw[i];
  FlatWalk wl(threadIdx.x, blockDim.x, n);
  for (int i = threadIdx.x; i < nt * n; i += blockDim.x, wl.next())
    lds[i] = in[(t0 + wl.q) * in_dist + wl.r];
  __syncthreads();
  lds_fft<kMask>(lds, nt, plan, twl, sgn);
  FlatWalk ws(threadIdx.x, blockDim.x, n);
  for (int i = threadIdx.x; i < nt * n; i += blockDim.x, ws.next())
    out[(t0 + ws.q) * out_dist + ws.r] = cscale(lds[ws.q * n + fft_pos(plan, ws.r)], scale);
}

__global__ void __launch_bounds__(kFftThreads)
r2c_batch_kernel(const float *__restrict__ in, long long in_dist, int in_len,
                 float2 *__restrict__ out, long long out_dist, FftPlanDev plan,
                 const float2 *__restrict__ tw, const float2 *__restrict__ fold, float scale) {
  extern __shared__ float2 lds[];
  const long long row = blockIdx.x;
  const float *x = in + row * in_dist;
  float2 *y = out + row * out_dist;
  const int M = plan.n;
  stage_real_tables(lds, plan, tw, fold);
  const float2 *twl = lds + M + 1;
  const float2 *fl = twl + tw_lds_count(M);
  real_pack(lds, M, [&](int i) { return i < in_len ? x[i] : 0.f; });
  lds_fft(lds, 1, plan, twl, -1.f);
  r2c_unfold(lds, plan, fl, [&](int k, float2 v) { y[k] = cscale(v, scale); });
}

__global__ void __launch_bounds__(kFftThreads)
c2r_batch_kernel(const float2 *__restrict__ in, long long in_dist, float *__restrict__ out,
                 long long out_dist, int out_len, FftPlanDev plan,
                 const float2 *__restrict__ tw, const float2 *__restrict__ fold, float scale) {
  extern __shared__ float2 lds[];
  const long long row = blockIdx.x;
  const float2 *X = in + row * in_dist;
  float *y = out + row * out_dist;
  const int M = plan.n;
  stage_real_tables(lds, plan, tw, fold);
  const float2 *twl = lds + M + 1;
  const float2 *fl = twl + tw_lds_count(M);
  c2r_fold(lds, M, fl, [&](int k) { return X[k]; });
  lds_fft(lds, 1, plan, twl, 1.f);
  c2r_unpack(lds, plan, [&](int i, float v) {
    if (i < out_len) y[i] = v * scale;
  });
}

int set_max_lds(const void *kernel, size_t bytes) {
  FXC_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return FXC_OK;
}

}  // namespace fxc

using namespace fxc;

extern "C" int fxc_fft_c2c(const float *d_in, long long in_dist, float *d_out, long long out_dist,
                           int n, long long batch, int inverse, float scale, void *stream) {
  FXC_CHECK_ARG(d_in && d_out && n > 0 && batch >= 0, "fxc_fft_c2c: invalid arguments");
  if (batch == 0) return FXC_OK;
  if (n + tw_lds_count(n) > kLdsComplexCapacity) {
    // beyond one workgroup's LDS: the four-step transform, per call scratch
    const LargeFft *lp = nullptr;
    FXC_TRY(large_plan(n, inverse ? 1 : -1, &lp));
    const long long chunk = std::min<long long>(batch, 65535);
    float2 *tmp = nullptr;
    FXC_TRY(dmalloc((void **)&tmp, (size_t)chunk * n * sizeof(float2)));
    int rc = FXC_OK;
    for (long long r = 0; r < batch && rc == FXC_OK; r += chunk)
      rc = large_c2c(lp, (const float2 *)d_in + r * in_dist, in_dist, (float2 *)d_out + r * out_dist,
                     out_dist, std::min(chunk, batch - r), tmp, scale, (hipStream_t)stream);
    if (rc == FXC_OK && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
      set_error("fxc_fft_c2c: stream synchronisation failed");
      rc = FXC_ERR_HIP;
    }
    dfree(tmp);
    return rc;
  }
  const FftPlan *p = nullptr;
  FXC_TRY(plan_cache_complex(n, inverse ? 1 : -1, &p));
  const bool smooth = plan_in_mask(p->dev, kRadixSmooth);
  const int nthr = smooth ? kFftThreads : kFftThreadsOddPrime;
  const int per = max_transforms_per_wg(p->dev, nthr, kLdsComplexCapacity);
  FXC_CHECK_ARG(per >= 1, "fxc_fft_c2c: length %d does not fit one workgroup", n);
  const size_t lds = ((size_t)per * n + tw_lds_count(n)) * sizeof(float2);
  const unsigned grid = (unsigned)ceil_div((size_t)batch, (size_t)per);
  if (smooth) {
    FXC_TRY(set_max_lds((const void *)c2c_batch_kernel<kRadixSmooth, kFftThreads>, lds));
    hipLaunchKernelGGL((c2c_batch_kernel<kRadixSmooth, kFftThreads>), dim3(grid), dim3(nthr), lds,
                       (hipStream_t)stream, (const float2 *)d_in, in_dist, (float2 *)d_out, out_dist,
                       batch, per, p->dev, p->d_tw, inverse ? 1.f : -1.f, scale);
  } else {
    FXC_TRY(set_max_lds((const void *)c2c_batch_kernel<kRadixAll, kFftThreadsOddPrime>, lds));
    hipLaunchKernelGGL((c2c_batch_kernel<kRadixAll, kFftThreadsOddPrime>), dim3(grid), dim3(nthr),
                       lds, (hipStream_t)stream, (const float2 *)d_in, in_dist, (float2 *)d_out,
                       out_dist, batch, per, p->dev, p->d_tw, inverse ? 1.f : -1.f, scale);
  }
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

extern "C" int fxc_fft_r2c(const float *d_in, long long in_dist, int in_len, float *d_out,
                           long long out_dist, int n, long long batch, float scale, void *stream) {
  FXC_CHECK_ARG(d_in && d_out && n >= 2 && (n % 2) == 0 && batch >= 0 && in_len >= 0,
                "fxc_fft_r2c: invalid arguments");
  if (batch == 0) return FXC_OK;
  if (real_lds_count(n / 2) > kLdsComplexCapacity) {
    const long long chunk = std::min<long long>(batch, 65535);
    float2 *work = nullptr;
    FXC_TRY(dmalloc((void **)&work, (size_t)chunk * large_real_work(n) * sizeof(float2)));
    int rc = FXC_OK;
    for (long long r = 0; r < batch && rc == FXC_OK; r += chunk)
      rc = large_r2c(n, d_in + r * in_dist, in_dist, in_len, (float2 *)d_out + r * out_dist,
                     out_dist, std::min(chunk, batch - r), work, scale, (hipStream_t)stream);
    if (rc == FXC_OK && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
      set_error("fxc_fft_r2c: stream synchronisation failed");
      rc = FXC_ERR_HIP;
    }
    dfree(work);
    return rc;
  }
  const RealFftPlan *p = nullptr;
  FXC_TRY(plan_cache_real(n, -1, &p));
  FXC_CHECK_ARG(max_transforms_per_wg(p->half.dev, kFftThreads, kLdsComplexCapacity) >= 1,
                "fxc_fft_r2c: length %d does not fit one workgroup", n);
  const size_t lds = (size_t)real_lds_count(n / 2) * sizeof(float2);
  FXC_CHECK_ARG(real_lds_count(n / 2) <= kLdsComplexCapacity, "fxc_fft_r2c: length %d too large", n);
  FXC_TRY(set_max_lds((const void *)r2c_batch_kernel, lds));
  hipLaunchKernelGGL(r2c_batch_kernel, dim3((unsigned)batch), dim3(kFftThreads), lds,
                     (hipStream_t)stream, d_in, in_dist, in_len, (float2 *)d_out, out_dist,
                     p->half.dev, p->half.d_tw, p->d_fold, scale);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

extern "C" int fxc_fft_c2r(const float *d_in, long long in_dist, float *d_out, long long out_dist,
                           int out_len, int n, long long batch, float scale, void *stream) {
  FXC_CHECK_ARG(d_in && d_out && n >= 2 && (n % 2) == 0 && batch >= 0 && out_len >= 0 &&
                    out_len <= n,
                "fxc_fft_c2r: invalid arguments");
  if (batch == 0) return FXC_OK;
  if (real_lds_count(n / 2) > kLdsComplexCapacity) {
    const long long chunk = std::min<long long>(batch, 65535);
    float2 *work = nullptr;
    FXC_TRY(dmalloc((void **)&work, (size_t)chunk * large_real_work(n) * sizeof(float2)));
    int rc = FXC_OK;
    for (long long r = 0; r < batch && rc == FXC_OK; r += chunk)
      rc = large_c2r(n, (const float2 *)d_in + r * in_dist, in_dist, d_out + r * out_dist,
                     out_dist, out_len, std::min(chunk, batch - r), work, (double)scale,
                     (hipStream_t)stream);
    if (rc == FXC_OK && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
      set_error("fxc_fft_c2r: stream synchronisation failed");
      rc = FXC_ERR_HIP;
    }
    dfree(work);
    return rc;
  }
  const RealFftPlan *p = nullptr;
  FXC_TRY(plan_cache_real(n, 1, &p));
  FXC_CHECK_ARG(max_transforms_per_wg(p->half.dev, kFftThreads, kLdsComplexCapacity) >= 1,
                "fxc_fft_c2r: length %d does not fit one workgroup", n);
  const size_t lds = (size_t)real_lds_count(n / 2) * sizeof(float2);
  FXC_CHECK_ARG(real_lds_count(n / 2) <= kLdsComplexCapacity, "fxc_fft_c2r: length %d too large", n);
  FXC_TRY(set_max_lds((const void *)c2r_batch_kernel, lds));
  hipLaunchKernelGGL(c2r_batch_kernel, dim3((unsigned)batch), dim3(kFftThreads), lds,
                     (hipStream_t)stream, (const float2 *)d_in, in_dist, d_out, out_dist, out_len,
                     p->half.dev, p->half.d_tw, p->d_fold, scale);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}
