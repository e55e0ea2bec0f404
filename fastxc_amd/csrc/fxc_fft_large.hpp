// Transforms longer than one workgroup's LDS: four-step FFT over HBM.
//
// The reference plans every length with cuFFT (native/sac2spec/device_memory.cu:23-32,
// native/xc/memory/memory.cu:478-483).  At dt 0.05 s its lengths are 72000
// (window), 72576 (prefilter) and 74088 (output / xc), whose half-length
// complex transforms (36000, 36288, 37044 points, 288-296 KB) exceed the 160 KB
// LDS of a gfx950 CU.  A length M = N1 N2 complex transform runs as
//   pass 1: N1-point transforms over the strided columns x[n1 N2 + n2], each
//           output multiplied by w_M^(n2 k1), stored Y[k1][n2];
//   pass 2: N2-point transforms over the rows Y[k1][.], stored X[k1 + N1 k2],
// both in LDS, C columns (rows) per workgroup so every global access is a
// run of C consecutive complex values.  Real transforms of length n = 2M use
// the packed half-length sequence z[j] = x[2j] + i x[2j+1] with a pack /
// unfold (R2C) or fold / unpack (C2R) pass.  Twiddles are FP64-computed
// tables.
#pragma once

#include "fxc_fft_device.hpp"

#include <hip/hip_runtime.h>

namespace fxc {

struct LargeFft {
  int M = 0, N1 = 0, N2 = 0, sign = -1;
  const FftPlan *p1 = nullptr;  // N1 points
  const FftPlan *p2 = nullptr;  // N2 points
  float2 *d_tw = nullptr;       // [M] exp(sign 2 pi i m / M)
  float2 *d_half = nullptr;     // [M + 1] exp(sign 2 pi i k / (2M)) (real fold / unfold)
};

// A factorisation M = N1 N2 with both factors within one workgroup's LDS
// transforms; false when none exists.
bool large_factor(int M, int *N1, int *N2);
// Cached plan of the complex length M (sign -1 forward, +1 inverse).
int large_plan(int M, int sign, const LargeFft **out);

// Complex transforms of `rows` rows, natural order in and out (unnormalised,
// times `scale`).  tmp: rows x M complex.  in may equal out.
int large_c2c(const LargeFft *p, const float2 *in, long long in_dist, float2 *out,
              long long out_dist, long long rows, float2 *tmp, float scale, hipStream_t st);
// Real forward transform of length n (n/2 = M): in_len samples per row (zero
// padded), bins 0..n/2 out.  work: 2 x rows x M complex.
int large_r2c(int n, const float *in, long long in_dist, int in_len, float2 *out,
              long long out_dist, long long rows, float2 *work, float scale, hipStream_t st);
// Real inverse transform of bins 0..n/2 (imaginary parts of 0 and n/2
// ignored), the first out_len samples out, each (float)(value * scale) with
// the product in double.  work: 2 x rows x M complex.
int large_c2r(int n, const float2 *in, long long in_dist, float *out, long long out_dist,
              int out_len, long long rows, float2 *work, double scale, hipStream_t st);
// Complex elements of work the real transforms of length n need per row.
inline long long large_real_work(int n) { return 2LL * (n / 2); }
// Frees every cached four-step plan's tables (fxc_release_plans).
void large_release();

// Fold of the real inverse: Z[k] = (X[k] + conj X[M-k]) + i w^k (X[k] - conj X[M-k]),
// k < M, w^k = half[k] of the inverse plan; grid (ceil(M/256), rows), 256 threads.
__global__ void large_fold_kernel(const float2 *__restrict__ X, long long in_dist, int M,
                                  const float2 *__restrict__ half, float2 *__restrict__ Z);

}  // namespace fxc
