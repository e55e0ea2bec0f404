// Four-step FFT over HBM for lengths beyond one workgroup's LDS (see
// fxc_fft_large.hpp for the decomposition).
#include "fxc_fft_large.hpp"
#include "fxc_plan_cache.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

namespace fxc {

constexpr int kLargeThreads = 512;
constexpr int kLargeCols = 64;  // columns (rows) per workgroup in each pass

// Pass 1: transforms of the columns n2 in [c0, c0 + C) of each row,
// Y[k1][n2] = w_M^(n2 k1) sum_n1 x[n1 N2 + n2] w_N1^(n1 k1).
template <unsigned kMask>
__global__ void __launch_bounds__(kLargeThreads)
large_cols_kernel(const float2 *__restrict__ in, long long in_dist, float2 *__restrict__ tmp,
                  int M, int N2, int C, FftPlanDev p1, const float2 *__restrict__ tw1,
                  const float2 *__restrict__ twM, float sgn) {
  extern __shared__ float2 lds[];
  const int N1 = p1.n;
  float2 *twl = lds + (size_t)C * N1;
  for (int i = threadIdx.x; i < kTwFine + p1.ncoarse; i += blockDim.x) twl[i] = tw1[i];
  const long long row = blockIdx.y;
  const int c0 = blockIdx.x * C;
  const int nc = min(C, N2 - c0);
  const float2 *x = in + row * in_dist;
  FlatWalk wl(threadIdx.x, blockDim.x, C);
  for (int i = threadIdx.x; i < C * N1; i += blockDim.x, wl.next()) {
    const int n1 = wl.q, c = wl.r;
    lds[c * N1 + n1] = c < nc ? x[(long long)n1 * N2 + c0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<kMask>(lds, C, p1, twl, sgn);
  float2 *y = tmp + row * (long long)M;
  FlatWalk ws(threadIdx.x, blockDim.x, nc);
  for (int i = threadIdx.x; i < nc * N1; i += blockDim.x, ws.next()) {
    const int k1 = ws.q, c = ws.r;
    const int n2 = c0 + c;
    const float2 v = lds[c * N1 + fft_pos(p1, k1)];
    // n2 < N2 and k1 < N1, so n2 k1 < M needs no reduction
    y[(long long)k1 * N2 + n2] = cmul(v, twM[n2 * k1]);
  }
}

// Pass 2: transforms of the rows k1 in [k0, k0 + C) of Y,
// X[k1 + N1 k2] = scale sum_n2 Y[k1][n2] w_N2^(n2 k2).
template <unsigned kMask>
__global__ void __launch_bounds__(kLargeThreads)
large_rows_kernel(const float2 *__restrict__ tmp, float2 *__restrict__ out, long long out_dist,
                  int M, int N1, int C, FftPlanDev p2, const float2 *__restrict__ tw2, float sgn,
                  float scale) {
  extern __shared__ float2 lds[];
  const int N2 = p2.n;
  float2 *twl = lds + (size_t)C * N2;
  for (int i = threadIdx.x; i < kTwFine + p2.ncoarse; i += blockDim.x) twl[i] = tw2[i];
  const long long row = blockIdx.y;
  const int k0 = blockIdx.x * C;
  const int nc = min(C, N1 - k0);
  const float2 *y = tmp + row * (long long)M;
  FlatWalk wl(threadIdx.x, blockDim.x, N2);
  for (int i = threadIdx.x; i < C * N2; i += blockDim.x, wl.next()) {
    const int c = wl.q, n2 = wl.r;
    lds[i] = c < nc ? y[(long long)(k0 + c) * N2 + n2] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<kMask>(lds, C, p2, twl, sgn);
  float2 *o = out + row * out_dist;
  FlatWalk ws(threadIdx.x, blockDim.x, nc);
  for (int i = threadIdx.x; i < nc * N2; i += blockDim.x, ws.next()) {
    const int k2 = ws.q, c = ws.r;
    o[k0 + c + (long long)N1 * k2] = cscale(lds[c * N2 + fft_pos(p2, k2)], scale);
  }
}

// z[j] = x[2j] + i x[2j+1] (zero past in_len).
__global__ void large_pack_kernel(const float *__restrict__ x, long long in_dist, int in_len,
                                  float2 *__restrict__ z, int M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const long long row = blockIdx.y;
  const float *xr = x + row * in_dist;
  const int a = 2 * j;
  z[row * M + j] = make_float2(a < in_len ? xr[a] : 0.f, a + 1 < in_len ? xr[a + 1] : 0.f);
}

// X[k] = (Z[k] + conj Z[M-k])/2 - (i/2) w^k (Z[k] - conj Z[M-k]), k = 0..M,
// w^k = exp(-2 pi i k / 2M).
__global__ void large_unfold_kernel(const float2 *__restrict__ Z, int M,
                                    const float2 *__restrict__ half, float2 *__restrict__ X,
                                    long long out_dist, float scale) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > M) return;
  const long long row = blockIdx.y;
  const float2 *zr = Z + row * M;
  const float2 a = zr[k == M ? 0 : k];
  const float2 b = cconj(zr[k == 0 ? 0 : M - k]);
  const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
  const float2 d = cmul(half[k], make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y)));
  // e - i d
  X[row * out_dist + k] = make_float2((e.x + d.y) * scale, (e.y - d.x) * scale);
}

// Z[k] = (X[k] + conj X[M-k]) + i w^k (X[k] - conj X[M-k]), k < M,
// w^k = exp(+2 pi i k / 2M); the imaginary parts of X[0] and X[M] are ignored.
__global__ void large_fold_kernel(const float2 *__restrict__ X, long long in_dist, int M,
                                  const float2 *__restrict__ half, float2 *__restrict__ Z) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const long long row = blockIdx.y;
  const float2 *xr = X + row * in_dist;
  float2 v;
  if (k == 0) {
    const float a = xr[0].x, b = xr[M].x;
    v = make_float2(a + b, a - b);
  } else {
    const float2 a = xr[k];
    const float2 b = cconj(xr[M - k]);
    const float2 s = cadd(a, b);
    const float2 d = cmul(half[k], csub(a, b));
    v = make_float2(s.x - d.y, s.y + d.x);
  }
  Z[row * M + k] = v;
}

// y[2j] = Re z[j], y[2j+1] = Im z[j], times scale (product in double), first
// out_len samples.
__global__ void large_unpack_kernel(const float2 *__restrict__ z, int M, float *__restrict__ y,
                                    long long out_dist, int out_len, double scale) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const long long row = blockIdx.y;
  const float2 v = z[row * M + j];
  float *yr = y + row * out_dist;
  if (2 * j < out_len) yr[2 * j] = (float)((double)v.x * scale);
  if (2 * j + 1 < out_len) yr[2 * j + 1] = (float)((double)v.y * scale);
}

namespace {
std::mutex g_large_mu;
std::map<std::tuple<int, int, int>, std::unique_ptr<LargeFft>> g_large;

// Columns per workgroup for an N-point sub-transform.
int large_cols(int N) {
  const int room = kLdsComplexCapacity - tw_lds_count(N) - 64;
  return std::max(1, std::min(kLargeCols, room / N));
}

dim3 grid_rows(long long n, long long rows, int threads) {
  return dim3((unsigned)ceil_div((size_t)n, (size_t)threads), (unsigned)rows);
}
}  // namespace

bool large_factor(int M, int *N1, int *N2) {
  int radix[kMaxPasses], np = 0;
  int best = 0;
  for (int d = 2; (long long)d * d <= M; ++d) {
    if (M % d) continue;
    const int e = M / d;
    if (e > 8192 || !factor_radices(d, radix, &np) || !factor_radices(e, radix, &np)) continue;
    best = d;  // the largest such d <= sqrt(M): the most balanced split
  }
  if (!best) return false;
  *N1 = best;
  *N2 = M / best;
  return true;
}

int large_plan(int M, int sign, const LargeFft **out) {
  int N1 = 0, N2 = 0;
  FXC_CHECK_ARG(M >= 4 && large_factor(M, &N1, &N2),
                "fft: length %d has no two-level factorisation into supported radices", M);
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  const FftPlan *p1 = nullptr, *p2 = nullptr;
  FXC_TRY(plan_cache_complex(N1, sign, &p1));
  FXC_TRY(plan_cache_complex(N2, sign, &p2));
  // the pass kernels compile the 2,3,5,7-smooth butterflies only
  FXC_CHECK_ARG(plan_in_mask(p1->dev, kRadixSmooth) && plan_in_mask(p2->dev, kRadixSmooth),
                "fft: length %d = %d x %d is not 2,3,5,7-smooth", M, N1, N2);
  std::lock_guard<std::mutex> lk(g_large_mu);
  auto key = std::make_tuple(dev, M, sign);
  auto it = g_large.find(key);
  if (it == g_large.end()) {
    std::unique_ptr<LargeFft> p(new LargeFft());
    p->M = M;
    p->N1 = N1;
    p->N2 = N2;
    p->sign = sign;
    p->p1 = p1;
    p->p2 = p2;
    std::vector<float2> tw((size_t)M), half((size_t)M + 1);
    for (int m = 0; m < M; ++m) {
      const double a = sign * 2.0 * M_PI * (double)m / (double)M;
      tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    for (int k = 0; k <= M; ++k) {
      const double a = sign * M_PI * (double)k / (double)M;
      half[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    half[M] = make_float2(-1.f, 0.f);  // exactly: the Nyquist bin stays real
    FXC_TRY(dmalloc((void **)&p->d_tw, tw.size() * sizeof(float2)));
    if (hipMemcpy(p->d_tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess ||
        dmalloc((void **)&p->d_half, half.size() * sizeof(float2)) != FXC_OK ||
        hipMemcpy(p->d_half, half.data(), half.size() * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess) {
      dfree(p->d_tw);
      dfree(p->d_half);
      set_error("fft: twiddle upload of length %d failed", M);
      return FXC_ERR_HIP;
    }
    it = g_large.emplace(key, std::move(p)).first;
  }
  *out = it->second.get();
  return FXC_OK;
}

int large_c2c(const LargeFft *p, const float2 *in, long long in_dist, float2 *out,
              long long out_dist, long long rows, float2 *tmp, float scale, hipStream_t st) {
  if (rows == 0) return FXC_OK;
  FXC_CHECK_ARG(rows <= 65535, "fft: at most 65535 rows per four-step call");
  const float sgn = (float)p->sign;
  const int C1 = large_cols(p->N1), C2 = large_cols(p->N2);
  const size_t lds1 = ((size_t)C1 * p->N1 + tw_lds_count(p->N1)) * sizeof(float2);
  const size_t lds2 = ((size_t)C2 * p->N2 + tw_lds_count(p->N2)) * sizeof(float2);
  auto k1 = large_cols_kernel<kRadixSmooth>;
  auto k2 = large_rows_kernel<kRadixSmooth>;
  FXC_HIP(hipFuncSetAttribute((const void *)k1, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds1));
  FXC_HIP(hipFuncSetAttribute((const void *)k2, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds2));
  hipLaunchKernelGGL(k1, dim3((unsigned)ceil_div(p->N2, C1), (unsigned)rows), dim3(kLargeThreads),
                     lds1, st, in, in_dist, tmp, p->M, p->N2, C1, p->p1->dev, p->p1->d_tw,
                     p->d_tw, sgn);
  FXC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k2, dim3((unsigned)ceil_div(p->N1, C2), (unsigned)rows), dim3(kLargeThreads),
                     lds2, st, tmp, out, out_dist, p->M, p->N1, C2, p->p2->dev, p->p2->d_tw, sgn,
                     scale);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

int large_r2c(int n, const float *in, long long in_dist, int in_len, float2 *out,
              long long out_dist, long long rows, float2 *work, float scale, hipStream_t st) {
  FXC_CHECK_ARG(n >= 8 && n % 2 == 0, "fft: real length %d must be even", n);
  const int M = n / 2;
  const LargeFft *p = nullptr;
  FXC_TRY(large_plan(M, -1, &p));
  float2 *w0 = work, *w1 = work + (size_t)rows * M;
  hipLaunchKernelGGL(large_pack_kernel, grid_rows(M, rows, 256), dim3(256), 0, st, in, in_dist,
                     in_len, w0, M);
  FXC_HIP(hipGetLastError());
  FXC_TRY(large_c2c(p, w0, M, w0, M, rows, w1, 1.f, st));
  hipLaunchKernelGGL(large_unfold_kernel, grid_rows(M + 1, rows, 256), dim3(256), 0, st, w0, M,
                     p->d_half, out, out_dist, scale);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

int large_c2r(int n, const float2 *in, long long in_dist, float *out, long long out_dist,
              int out_len, long long rows, float2 *work, double scale, hipStream_t st) {
  FXC_CHECK_ARG(n >= 8 && n % 2 == 0, "fft: real length %d must be even", n);
  const int M = n / 2;
  const LargeFft *p = nullptr;
  FXC_TRY(large_plan(M, 1, &p));
  float2 *w0 = work, *w1 = work + (size_t)rows * M;
  hipLaunchKernelGGL(large_fold_kernel, grid_rows(M, rows, 256), dim3(256), 0, st, in, in_dist, M,
                     p->d_half, w0);
  FXC_HIP(hipGetLastError());
  FXC_TRY(large_c2c(p, w0, M, w0, M, rows, w1, 1.f, st));
  hipLaunchKernelGGL(large_unpack_kernel, grid_rows(M, rows, 256), dim3(256), 0, st, w0, M, out,
                     out_dist, out_len, scale);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

void large_release() {
  std::lock_guard<std::mutex> lk(g_large_mu);
  for (auto &kv : g_large) {
    dfree(kv.second->d_tw);
    dfree(kv.second->d_half);
  }
  g_large.clear();
}

}  // namespace fxc
