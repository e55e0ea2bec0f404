// Process-wide cache of FFT plans keyed by (device, length, sign, radix
// options).  Plans own
// small device twiddle tables; they live until fxc_release_plans() or exit.
#pragma once

#include "fxc_fft_real.hpp"

namespace fxc {

// Complex values one workgroup can stage in LDS (160 KB on gfx950),
// including the FFT twiddle table.
constexpr int kLdsComplexCapacity = 20480;

// max_radix bounds the butterfly size (part of the cache key).  With
// create = false a plan that is not cached yet is an argument error
// instead of being built (no device allocation).
int plan_cache_complex(int n, int sign, const FftPlan **out, int max_radix = 32,
                       bool create = true);
int plan_cache_real(int n, int sign, const RealFftPlan **out, int max_radix = 32,
                    bool largest_last = false);
void plan_cache_release();

// Bluestein (chirp-z) form of a length-n DFT for lengths with a prime factor
// no butterfly covers:
//   X_s[k] = c_s(k) sum_{j<n} x[j] c_s(j) conj(c_s(k-j)),  c_s(j) = exp(s i pi j^2 / n),
// evaluated as a circular convolution of length m >= 2n-1 (m factorable):
// forward m-point DIF transform, product with B_s = DFT_m(conj(c_s)) / m
// held in the same digit-reversed layout, m-point DIT inverse.
struct BluesteinPlan {
  int n = 0, m = 0;
  const FftPlan *fwd = nullptr;   // m points, sign -1
  const FftPlan *inv = nullptr;   // m points, sign +1
  float2 *d_chirp = nullptr;      // [n] c_{-1}(j) = exp(-i pi j^2 / n), FP64 phases
  float2 *d_bhat_f = nullptr;     // [m] B_{-1}, digit-reversed, / m
  float2 *d_bhat_i = nullptr;     // [m] B_{+1}
};
// Smallest m >= 2n-1 that factor_radices accepts (0 when none up to 2^16).
int bluestein_length(int n);
int plan_cache_bluestein(int n, const BluesteinPlan **out, bool create = true);

}  // namespace fxc
