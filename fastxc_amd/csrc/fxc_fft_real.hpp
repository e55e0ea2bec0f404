// Real-to-complex and complex-to-real transforms built on the workgroup FFT.
//
// A real sequence x of even length N is packed as z[j] = x[2j] + i x[2j+1]
// and transformed with one complex FFT of length M = N/2; a fold with the
// table w_k = exp(sign*2*pi*i*k/N) converts between Z and the one-sided
// spectrum X[0..M].  All load and store steps take functors so the callers
// fuse their element-wise pre/post-processing into the FFT passes.
#pragma once

#include "fxc_fft_device.hpp"

namespace fxc {

struct RealFftPlan {
  int n = 0;            // real length (even)
  FftPlan half;         // complex plan of length n/2
  // exp(sign*2*pi*i*k/n), k = 0..n/2, as a two-level table
  // [fine(128) | coarse(ceil((n/2+1)/128))], staged in LDS by the kernels
  float2 *d_fold = nullptr;
  int fold_count = 0;
};

// LDS layout of a single real transform of length n = 2M:
// [data: M+1 complex][twiddles of the M-point plan][fold table].
inline __host__ __device__ int fold_lds_count(int M) { return kTwFine + (M + 1 + kTwFine - 1) / kTwFine; }
inline __host__ __device__ int real_lds_count(int M) { return M + 1 + tw_lds_count(M) + fold_lds_count(M); }

__device__ __forceinline__ float2 fold_lookup(const float2 *fold, int k) {
  return cmul(fold[kTwFine + (k >> 7)], fold[k & (kTwFine - 1)]);
}

// Copies the twiddle and fold tables into their LDS slots (after M+1 data
// slots) and synchronises the workgroup.
__device__ inline void stage_real_tables(float2 *lds, const FftPlanDev &plan,
                                         const float2 *__restrict__ tw,
                                         const float2 *__restrict__ fold) {
  const int M = plan.n;
  float2 *twl = lds + M + 1;
  float2 *fl = twl + tw_lds_count(M);
  for (int i = threadIdx.x; i < kTwFine + plan.ncoarse; i += blockDim.x) twl[i] = tw[i];
  for (int i = threadIdx.x; i < fold_lds_count(M); i += blockDim.x) fl[i] = fold[i];
  __syncthreads();
}

int real_fft_plan_create(RealFftPlan *plan, int n, int sign, int max_radix = 32,
                         bool largest_last = false);
void real_fft_plan_destroy(RealFftPlan *plan);
// Largest number of transforms one 1024-thread workgroup can hold.
int max_transforms_per_wg(const FftPlanDev &plan, int nthreads, int lds_complex_capacity);

// lds[j] = ld(2j) + i*ld(2j+1), j < M.
template <class LoadReal>
__device__ inline void real_pack(float2 *lds, int M, const LoadReal &ld) {
#pragma unroll 8
  for (int j = threadIdx.x; j < M; j += blockDim.x) lds[j] = make_float2(ld(2 * j), ld(2 * j + 1));
  __syncthreads();
}

// From Z = FFT_M(z) (forward, digit-reversed in LDS) emit the one-sided
// spectrum X[k], k = 0..M:
//   X[k] = E_k + w_k O_k, E_k = (Z[k] + conj Z[M-k])/2, O_k = -i (Z[k] - conj Z[M-k])/2,
// with w_k = exp(-2*pi*i*k/N) and Z[M] = Z[0].
template <class StoreSpec>
__device__ inline void r2c_unfold(const float2 *lds, const FftPlanDev &plan, const float2 *fold,
                                  const StoreSpec &st) {
  const int M = plan.n;
  for (int k = threadIdx.x; k <= M; k += blockDim.x) {
    const float2 a = lds[fft_pos(plan, k == M ? 0 : k)];
    const float2 b = cconj(lds[fft_pos(plan, k == 0 ? 0 : M - k)]);
    const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
    const float2 d = csub(a, b);
    const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
    float2 x = cadd(e, cmul(fold_lookup(fold, k), o));
    if (k == 0 || k == M) x.y = 0.f;
    st(k, x);
  }
}

// Builds Z[k] = (X[k] + conj X[M-k]) + i w_k (X[k] - conj X[M-k]) in LDS for
// an inverse FFT of length M (w_k = exp(+2*pi*i*k/N)).  The imaginary parts
// of X[0] and X[M] are ignored, as a real inverse transform requires.
template <class LoadSpec>
__device__ inline void c2r_fold(float2 *lds, int M, const float2 *fold, const LoadSpec &ld) {
#pragma unroll 8
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    float2 a = ld(k);
    float2 b = cconj(ld(M - k));
    if (k == 0) {
      a.y = 0.f;
      b.y = 0.f;
    }
    const float2 s = cadd(a, b);
    const float2 d = cmul(fold_lookup(fold, k), csub(a, b));
    lds[k] = make_float2(s.x - d.y, s.y + d.x);
  }
  __syncthreads();
}

// x[2j] = Re z[j], x[2j+1] = Im z[j] (z digit-reversed in LDS).
template <class StoreReal>
__device__ inline void c2r_unpack(const float2 *lds, const FftPlanDev &plan, const StoreReal &st) {
  const int M = plan.n;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    const float2 z = lds[fft_pos(plan, j)];
    st(2 * j, z.x);
    st(2 * j + 1, z.y);
  }
}

}  // namespace fxc
