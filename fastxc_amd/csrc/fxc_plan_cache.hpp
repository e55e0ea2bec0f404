// Process-wide cache of FFT plans keyed by (device, length, sign, radix
// options).  Plans own
// small device twiddle tables; they live until fxc_release_plans() or exit.
#pragma once

#include "fxc_fft_real.hpp"

namespace fxc {

// Complex values one workgroup can stage in LDS (160 KB on gfx950),
// including the FFT twiddle table.
constexpr int kLdsComplexCapacity = 20480;

// max_radix bounds the butterfly size (part of the cache key).
int plan_cache_complex(int n, int sign, const FftPlan **out, int max_radix = 32);
int plan_cache_real(int n, int sign, const RealFftPlan **out, int max_radix = 32,
                    bool largest_last = false);
void plan_cache_release();

}  // namespace fxc
