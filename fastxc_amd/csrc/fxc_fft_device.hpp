// Workgroup-level mixed-radix FFT staged in LDS (device code).
//
// One workgroup owns `ntrans` transforms of length n laid out back to back in
// LDS.  Passes are in-place decimation-in-frequency butterflies, so a thread
// holds one butterfly at a time and a single LDS buffer suffices: a
// 37044-point real transform (an 18522-point complex one, 148 KB) lives in
// the 160 KB LDS of one CU.  The output stays in digit-reversed order and
// every consumer reads it through fft_pos(), which the fused kernels need
// anyway because they only read the bins or lags they keep.
//
// Butterflies are written for radices 2, 3, 4, 5, 7, 8, the composites 6, 9,
// 10, 12, 14, 15, 16, 21 (two register-resident stages each) and the odd
// primes the correlation lengths need (23, 29); all constants are computed at
// compile time in double precision and rounded once to float.
#pragma once

#include "fxc_common.hpp"

namespace fxc {

__host__ __device__ constexpr double cx_pi() { return 3.14159265358979323846264338327950288; }

// Taylor series on [-pi, pi]; 40 terms give full double precision there.
__host__ __device__ constexpr double cx_sin_reduced(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 40; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
__host__ __device__ constexpr double cx_cos_reduced(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 40; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// cos/sin of 2*pi*j/r for integer j, r (argument reduced exactly).
__host__ __device__ constexpr double cx_cos2pi(int j, int r) {
  int jj = ((j % r) + r) % r;
  if (2 * jj > r) jj -= r;
  return cx_cos_reduced(2.0 * cx_pi() * jj / r);
}
__host__ __device__ constexpr double cx_sin2pi(int j, int r) {
  int jj = ((j % r) + r) % r;
  if (2 * jj > r) jj -= r;
  return cx_sin_reduced(2.0 * cx_pi() * jj / r);
}

template <int R>
struct RootTable {
  float c[R];
  float s[R];
  constexpr RootTable() : c(), s() {
    for (int j = 0; j < R; ++j) {
      c[j] = (float)cx_cos2pi(j, R);
      s[j] = (float)cx_sin2pi(j, R);
    }
  }
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// i * sgn * a
__device__ __forceinline__ float2 cmul_isgn(float2 a, float sgn) { return make_float2(-sgn * a.y, sgn * a.x); }

// In-place DFT of an odd number R of values:
//   b_q = sum_m a_m exp(sgn*2*pi*i*m*q/R).
// Pairs m with R-m so each pair costs one real multiply per constant:
//   b_q = a_0 + sum_m [t_m cos(2*pi*m*q/R) + i*sgn*u_m sin(2*pi*m*q/R)],
// with t_m = a_m + a_{R-m} and u_m = a_m - a_{R-m}.
template <int R>
__device__ __forceinline__ void dft_generic(float2 *a, float sgn) {
  static_assert((R & 1) == 1, "generic butterfly is for odd radices");
  constexpr RootTable<R> tb{};
  constexpr int H = (R - 1) / 2;
  float2 t[H + 1], u[H + 1];
#pragma unroll
  for (int m = 1; m <= H; ++m) {
    t[m] = cadd(a[m], a[R - m]);
    u[m] = csub(a[m], a[R - m]);
  }
  float2 b[R];
  float2 s0 = a[0];
#pragma unroll
  for (int m = 1; m <= H; ++m) s0 = cadd(s0, t[m]);
  b[0] = s0;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    float2 re = a[0];
    float2 im = make_float2(0.f, 0.f);
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      const int idx = (m * q) % R;
      re.x += t[m].x * tb.c[idx];
      re.y += t[m].y * tb.c[idx];
      im.x += u[m].x * tb.s[idx];
      im.y += u[m].y * tb.s[idx];
    }
    const float2 iim = cmul_isgn(im, sgn);
    b[q] = cadd(re, iim);
    b[R - q] = csub(re, iim);
  }
#pragma unroll
  for (int q = 0; q < R; ++q) a[q] = b[q];
}

template <int R>
__device__ __forceinline__ void dft(float2 *a, float sgn) {
  dft_generic<R>(a, sgn);
}

// The odd-radix DFT of dft_generic with its outputs handed to emit(q, b_q)
// as they are formed (q = 0, then the pairs q, R-q), so no output array is
// live: a[m] and a[R-m] are overwritten by t_m and u_m.  Used by the passes
// for the large prime radices, where an R-value output array would double
// the registers the butterfly holds.
template <int R, typename Emit>
__device__ __forceinline__ void dft_odd_stream(float2 *a, float sgn, Emit emit) {
  static_assert((R & 1) == 1, "streaming butterfly is for odd radices");
  constexpr RootTable<R> tb{};
  constexpr int H = (R - 1) / 2;
  float2 s0 = a[0];
#pragma unroll
  for (int m = 1; m <= H; ++m) {
    const float2 x = a[m], y = a[R - m];
    a[m] = cadd(x, y);
    a[R - m] = csub(x, y);
    s0 = cadd(s0, a[m]);
  }
  emit(0, s0);
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    float2 re = a[0];
    float2 im = make_float2(0.f, 0.f);
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      const int idx = (m * q) % R;
      re.x += a[m].x * tb.c[idx];
      re.y += a[m].y * tb.c[idx];
      im.x += a[R - m].x * tb.s[idx];
      im.y += a[R - m].y * tb.s[idx];
    }
    const float2 iim = cmul_isgn(im, sgn);
    emit(q, cadd(re, iim));
    emit(R - q, csub(re, iim));
    // one output pair in flight: keeps the accumulators of later pairs
    // from being hoisted (registers, not ILP, bound these butterflies)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Radices whose passes use dft_odd_stream.
template <int R>
constexpr bool kStreamRadix = (R == 23 || R == 29);

template <>
__device__ __forceinline__ void dft<2>(float2 *a, float) {
  const float2 x0 = a[0], x1 = a[1];
  a[0] = cadd(x0, x1);
  a[1] = csub(x0, x1);
}

template <>
__device__ __forceinline__ void dft<4>(float2 *a, float sgn) {
  const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
  const float2 s13 = cadd(a[1], a[3]), d13 = cmul_isgn(csub(a[1], a[3]), sgn);
  a[0] = cadd(s02, s13);
  a[2] = csub(s02, s13);
  a[1] = cadd(d02, d13);
  a[3] = csub(d02, d13);
}

template <>
__device__ __forceinline__ void dft<8>(float2 *a, float sgn) {
  // 8 = 2 x 4: two radix-4 DFTs on even/odd samples, twiddle, radix-2.
  float2 e[4] = {a[0], a[2], a[4], a[6]};
  float2 o[4] = {a[1], a[3], a[5], a[7]};
  dft<4>(e, sgn);
  dft<4>(o, sgn);
  constexpr float r2 = 0.70710678118654752440f;
  // o[q] *= exp(sgn*2*pi*i*q/8)
  o[1] = make_float2(r2 * (o[1].x - sgn * o[1].y), r2 * (o[1].y + sgn * o[1].x));
  o[2] = cmul_isgn(o[2], sgn);
  o[3] = make_float2(r2 * (-o[3].x - sgn * o[3].y), r2 * (-o[3].y + sgn * o[3].x));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a[q] = cadd(e[q], o[q]);
    a[q + 4] = csub(e[q], o[q]);
  }
}

// Composite DFT of R = R1*R2 values in registers (one Cooley-Tukey split:
// R2 DFTs of length R1, internal twiddles exp(sgn*2*pi*i*n2*k1/R) from a
// compile-time table, R1 DFTs of length R2):
//   X[k1 + R1*k2] = sum_n2 W_R2^(n2 k2) W_R^(n2 k1) sum_n1 a[R2 n1 + n2] W_R1^(n1 k1).
template <int R1, int R2>
__device__ __forceinline__ void dft_composite(float2 *a, float sgn) {
  constexpr int R = R1 * R2;
  constexpr RootTable<R> tb{};
  float2 t[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    float2 x[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) x[n1] = a[R2 * n1 + n2];
    dft<R1>(x, sgn);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int idx = (n2 * k1) % R;
      if (idx != 0) x[k1] = cmul(x[k1], make_float2(tb.c[idx], sgn * tb.s[idx]));
      t[n2][k1] = x[k1];
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    float2 y[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) y[n2] = t[n2][k1];
    dft<R2>(y, sgn);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = y[k2];
  }
}

template <> __device__ __forceinline__ void dft<6>(float2 *a, float sgn) { dft_composite<2, 3>(a, sgn); }
template <> __device__ __forceinline__ void dft<9>(float2 *a, float sgn) { dft_composite<3, 3>(a, sgn); }
template <> __device__ __forceinline__ void dft<10>(float2 *a, float sgn) { dft_composite<2, 5>(a, sgn); }
template <> __device__ __forceinline__ void dft<12>(float2 *a, float sgn) { dft_composite<4, 3>(a, sgn); }
template <> __device__ __forceinline__ void dft<14>(float2 *a, float sgn) { dft_composite<2, 7>(a, sgn); }
template <> __device__ __forceinline__ void dft<15>(float2 *a, float sgn) { dft_composite<3, 5>(a, sgn); }
template <> __device__ __forceinline__ void dft<16>(float2 *a, float sgn) { dft_composite<4, 4>(a, sgn); }
template <> __device__ __forceinline__ void dft<21>(float2 *a, float sgn) { dft_composite<3, 7>(a, sgn); }

// Largest butterfly count per thread a plan may need (host-side check only;
// threads loop over their butterflies one at a time).
constexpr int kMaxElemsPerThread = 1 << 20;

__device__ __forceinline__ unsigned udiv_magic(unsigned x, unsigned d, unsigned magic) {
  return d <= 1u ? x : __umulhi(x, magic);
}

// w_n^e from the LDS two-level table (tw = [fine(128) | coarse]).
__device__ __forceinline__ float2 tw_lookup(const float2 *tw, int e) {
  return cmul(tw[kTwFine + (e >> 7)], tw[e & (kTwFine - 1)]);
}

// Multiplies v[1..R-1] by w^(q*e), q = 1..R-1 (exponents modulo n).  Up to
// radix 16 the powers come from one table read and a chain of products
// (at most 14 roundings, ~1e-6 relative); larger radices read each power.
template <int R>
__device__ __forceinline__ void apply_twiddles(float2 *v, const float2 *tw, int e, int n) {
  if constexpr (R <= 16) {
    const float2 w = tw_lookup(tw, e);
    float2 wq = w;
    v[1] = cmul(v[1], w);
#pragma unroll
    for (int q = 2; q < R; ++q) {
      wq = cmul(wq, w);
      v[q] = cmul(v[q], wq);
    }
  } else {
    int eq = 0;
#pragma unroll
    for (int q = 1; q < R; ++q) {
      eq += e;
      if (eq >= n) eq -= n;
      v[q] = cmul(v[q], tw_lookup(tw, eq));
    }
  }
}

// Quotient q and remainder r of i = q d + r along i = i0, i0 + step, ...:
// one division at the start, then an add and a compare per step (a runtime
// integer division per element costs a few dozen VALU instructions).
struct FlatWalk {
  int q, r, dq, dr, d;
  __device__ __forceinline__ FlatWalk(int i0, int step, int d_) : d(d_) {
    q = i0 / d;
    r = i0 - q * d;
    dq = step / d;
    dr = step - dq * d;
  }
  __device__ __forceinline__ void next() {
    q += dq;
    r += dr;
    if (r >= d) {
      r -= d;
      ++q;
    }
  }
};

// One in-place DIF pass (butterfly, then twiddle) over ntrans transforms.
template <int R>
__device__ __forceinline__ void dif_pass(float2 *buf, int ntrans, const FftPlanDev &plan, int p,
                                         const float2 *tw, float sgn, int tid, int nthr) {
  const int n = plan.n;
  const int span = plan.span[p];
  const int stride = plan.stride[p];
  const int tstep = plan.tstep[p];
  const unsigned smagic = plan.smagic[p];
  const unsigned bmagic = plan.bmagic[p];
  const int nbt = n / R;
  const int total = ntrans * nbt;
  for (int b = tid; b < total; b += nthr) {
    const int t = ntrans > 1 ? (int)udiv_magic((unsigned)b, (unsigned)nbt, bmagic) : 0;
    const int r = b - t * nbt;
    const int blk = (int)udiv_magic((unsigned)r, (unsigned)stride, smagic);
    const int j = r - blk * stride;
    float2 *base = buf + t * n + blk * span + j;
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = base[m * stride];
    if constexpr (kStreamRadix<R>) {
      // outputs leave as formed: b_q x w^(q e), e = j tstep (mod n).  The
      // output offset goes through an empty asm operand so the compiler
      // re-derives the store addresses instead of holding R load addresses.
      int off = t * n + blk * span + j;
      asm volatile("" : "+v"(off));
      float2 *obase = buf + off;
      const int e = j * tstep;
      // j < stride = span / R and tstep = n / span, so e < n / R: R e needs
      // no reduction modulo n
      const int eR = e * R;
      int eq = 0;
      dft_odd_stream<R>(v, sgn, [&](int q, float2 y) {
        if (q != 0 && j > 0) {
          int x;
          if (q <= (R - 1) / 2) {
            eq += e;
            if (eq >= n) eq -= n;
            x = eq;
          } else {
            x = eR - eq;  // (R - q) e = R e - q e, q e being the last eq
            if (x < 0) x += n;
          }
          y = cmul(y, tw_lookup(tw, x));
        }
        obase[q * stride] = y;
      });
    } else {
      dft<R>(v, sgn);
      if (j > 0) apply_twiddles<R>(v, tw, j * tstep, n);
      float2 *obase = base;
      if constexpr (R >= 12) {  // as above, for the large composite butterflies
        int off = t * n + blk * span + j;
        asm volatile("" : "+v"(off));
        obase = buf + off;
      }
#pragma unroll
      for (int q = 0; q < R; ++q) obase[q * stride] = v[q];
    }
  }
  __syncthreads();
}

// One in-place DIT pass (twiddle, then butterfly): the exact inverse
// ordering of dif_pass, consuming digit-reversed data.
template <int R>
__device__ __forceinline__ void dit_pass(float2 *buf, int ntrans, const FftPlanDev &plan, int p,
                                         const float2 *tw, float sgn, int tid, int nthr) {
  const int n = plan.n;
  const int span = plan.span[p];
  const int stride = plan.stride[p];
  const int tstep = plan.tstep[p];
  const unsigned smagic = plan.smagic[p];
  const unsigned bmagic = plan.bmagic[p];
  const int nbt = n / R;
  const int total = ntrans * nbt;
  for (int b = tid; b < total; b += nthr) {
    const int t = ntrans > 1 ? (int)udiv_magic((unsigned)b, (unsigned)nbt, bmagic) : 0;
    const int r = b - t * nbt;
    const int blk = (int)udiv_magic((unsigned)r, (unsigned)stride, smagic);
    const int j = r - blk * stride;
    float2 *base = buf + t * n + blk * span + j;
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = base[m * stride];
    if (j > 0) apply_twiddles<R>(v, tw, j * tstep, n);
    if constexpr (kStreamRadix<R>) {
      int off = t * n + blk * span + j;  // see dif_pass
      asm volatile("" : "+v"(off));
      float2 *obase = buf + off;
      dft_odd_stream<R>(v, sgn, [&](int q, float2 y) { obase[q * stride] = y; });
    } else {
      dft<R>(v, sgn);
      float2 *obase = base;
      if constexpr (R >= 12) {  // see dif_pass
        int off = t * n + blk * span + j;
        asm volatile("" : "+v"(off));
        obase = buf + off;
      }
#pragma unroll
      for (int q = 0; q < R; ++q) obase[q * stride] = v[q];
    }
  }
  __syncthreads();
}

// Position of bin k in the digit-reversed LDS result.
__device__ __forceinline__ int fft_pos(const FftPlanDev &plan, int k) {
  int pos = 0;
  for (int p = 0; p < plan.npass; ++p) {
    const int r = plan.radix[p];
    const int q = (int)udiv_magic((unsigned)k, (unsigned)r, plan.rmagic[p]);
    pos += (k - q * r) * plan.stride[p];
    k = q;
  }
  return pos;
}

// Copies the plan's twiddle table (global) into LDS; ends with a barrier.
__device__ inline void stage_twiddles(float2 *tw_lds, const FftPlanDev &plan,
                                      const float2 *__restrict__ tw_global) {
  for (int i = threadIdx.x; i < kTwFine + plan.ncoarse; i += blockDim.x) tw_lds[i] = tw_global[i];
  __syncthreads();
}

// Radix sets a kernel compiles butterflies for (bit r = radix r).  Kernels on
// 2,3,5,7-smooth lengths leave out the large odd primes so 1024-thread
// workgroups stay within 128 VGPRs.
constexpr unsigned kRadixSmooth = (1u << 2) | (1u << 3) | (1u << 4) | (1u << 5) | (1u << 6) |
                                  (1u << 7) | (1u << 8) | (1u << 9) | (1u << 10) | (1u << 12) |
                                  (1u << 14) | (1u << 15) | (1u << 16) | (1u << 21);
constexpr unsigned kRadixAll = kRadixSmooth | (1u << 23) | (1u << 29);

#define FXC_RADIX_SWITCH(FN)                                                  \
  switch (plan.radix[p]) {                                                    \
    case 2: if constexpr ((kMask >> 2) & 1u) FN<2>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 3: if constexpr ((kMask >> 3) & 1u) FN<3>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 4: if constexpr ((kMask >> 4) & 1u) FN<4>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 5: if constexpr ((kMask >> 5) & 1u) FN<5>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 7: if constexpr ((kMask >> 7) & 1u) FN<7>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 8: if constexpr ((kMask >> 8) & 1u) FN<8>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 6: if constexpr ((kMask >> 6) & 1u) FN<6>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 9: if constexpr ((kMask >> 9) & 1u) FN<9>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break;   \
    case 10: if constexpr ((kMask >> 10) & 1u) FN<10>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 12: if constexpr ((kMask >> 12) & 1u) FN<12>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 14: if constexpr ((kMask >> 14) & 1u) FN<14>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 15: if constexpr ((kMask >> 15) & 1u) FN<15>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 16: if constexpr ((kMask >> 16) & 1u) FN<16>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 21: if constexpr ((kMask >> 21) & 1u) FN<21>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 23: if constexpr ((kMask >> 23) & 1u) FN<23>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    case 29: if constexpr ((kMask >> 29) & 1u) FN<29>(buf, ntrans, plan, p, tw, sgn, tid, nthr); break; \
    default: break;                                                           \
  }

// Natural-order input -> digit-reversed output (tw: LDS twiddle table).
template <unsigned kMask = kRadixSmooth>
__device__ __forceinline__ void lds_fft(float2 *buf, int ntrans, const FftPlanDev &plan, const float2 *tw,
                               float sgn) {
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  for (int p = 0; p < plan.npass; ++p) {
    FXC_RADIX_SWITCH(dif_pass)
  }
}

// The first npass_run DIF passes only.  After passes 0..P-2 every block of
// span R = radix[P-1] contiguous values holds the input of an R-point DFT
// whose output t is bin k = lower + (n/R) t, where the block starts at the
// digit-reversed position of lower = k mod (n/R) (see last_pass_bin).
template <unsigned kMask = kRadixSmooth>
__device__ __forceinline__ void lds_fft_head(float2 *buf, int ntrans, const FftPlanDev &plan,
                                    const float2 *tw, float sgn, int npass_run) {
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  for (int p = 0; p < npass_run; ++p) {
    FXC_RADIX_SWITCH(dif_pass)
  }
}

// Bin k of a transform whose last pass was left out (lds_fft_head with
// npass_run = P-1): the direct R-point DFT of its block,
//   Z[k] = sum_{m<R} buf[base + m] w_R^(m t),  w_R = w_n^(n/R),
// with the powers w_R^(m t) re-read from the table every 8 terms.
__device__ inline float2 last_pass_bin(const float2 *buf, const FftPlanDev &plan,
                                       const float2 *tw, int k) {
  const int P = plan.npass;
  if (P == 0) return buf[0];
  const int R = plan.radix[P - 1];
  const int g = plan.n / R;
  const int t = (int)udiv_magic((unsigned)k, (unsigned)g, plan.bmagic[P - 1]);
  int lower = k - t * g;
  int base = 0;
  for (int p = 0; p < P - 1; ++p) {
    const int r = plan.radix[p];
    const int q = (int)udiv_magic((unsigned)lower, (unsigned)r, plan.rmagic[p]);
    base += (lower - q * r) * plan.stride[p];
    lower = q;
  }
  const float2 wt = tw_lookup(tw, t * g);
  float2 z = buf[base];
  float2 wm = wt;
  int e = t;  // (m t) mod R for m = 1
  for (int m = 1; m < R; ++m) {
    const float2 v = buf[base + m];
    z.x += v.x * wm.x - v.y * wm.y;
    z.y += v.x * wm.y + v.y * wm.x;
    e += t;
    if (e >= R) e -= R;
    wm = ((m + 1) & 7) == 0 ? tw_lookup(tw, e * g) : cmul(wm, wt);
  }
  return z;
}

// Digit-reversed input -> natural-order output (the inverse ordering of
// lds_fft with the same plan radices; tw built for the sign sgn).
template <unsigned kMask = kRadixSmooth>
__device__ __forceinline__ void lds_fft_dit(float2 *buf, int ntrans, const FftPlanDev &plan,
                                   const float2 *tw, float sgn) {
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  for (int p = plan.npass - 1; p >= 0; --p) {
    FXC_RADIX_SWITCH(dit_pass)
  }
}

// True when every radix of `plan` is in kMask (host check before launch).
inline bool plan_in_mask(const FftPlanDev &plan, unsigned mask) {
  for (int p = 0; p < plan.npass; ++p)
    if (!((mask >> plan.radix[p]) & 1u)) return false;
  return true;
}

}  // namespace fxc
