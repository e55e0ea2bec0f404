// Workgroup FFT with a compile-time plan (device code + host table builder).
//
// The runtime-plan passes of fxc_fft_device.hpp spend a large share of their
// VALU work on index arithmetic (magic-number divisions by runtime strides)
// and on the two-level twiddle lookup.  For the one length the correlation
// hot path runs at production size, the plan is a template argument instead:
// strides, divisions and LDS offsets fold into constants.  The first pass
// (the one with the most distinct twiddles) reads w^(j q), q = 1..R-1,
// ready-made from an L2-resident global table; later passes read w^j from a
// small LDS table and chain its powers.
#pragma once

#include "fxc_fft_device.hpp"

#include <cmath>
#include <vector>

namespace fxc {

// DIF plan: radices in pass order; pass p works on spans of span(p) values
// with stride(p) = span(p) / R_p between the legs of a butterfly.
template <int M, int... Rs>
struct FixedPlan {
  static constexpr int kM = M;
  static constexpr int kPasses = (int)sizeof...(Rs);
  static constexpr int kRadix[kPasses] = {Rs...};
  static constexpr int span(int p) {
    int s = M;
    for (int i = 0; i < p; ++i) s /= kRadix[i];
    return s;
  }
  static constexpr int stride(int p) { return span(p) / kRadix[p]; }
  // twiddle row of one butterfly: R-1 values padded to an even count so
  // every row starts on a 16-byte boundary
  static constexpr int tw_pitch(int p) { return (kRadix[p] - 1 + 1) / 2 * 2; }
  static constexpr int tw_offset(int p) {
    int o = 0;
    for (int i = 0; i < p; ++i) o += stride(i) * tw_pitch(i);
    return o;
  }
  static constexpr bool valid() {
    int prod = 1;
    for (int i = 0; i < kPasses; ++i) prod *= kRadix[i];
    return prod == M;
  }
};

// Host: the twiddle table of passes 0..npass-1 for transform sign `sign`:
// row j of pass p holds exp(sign*2*pi*i*j*q/span(p)), q = 1..R-1.
template <class FP>
std::vector<float2> fixed_twiddles(int npass, int sign) {
  std::vector<float2> t((size_t)FP::tw_offset(npass), make_float2(0.f, 0.f));
  for (int p = 0; p < npass; ++p) {
    const int R = FP::kRadix[p], S = FP::span(p), s = FP::stride(p);
    for (int j = 0; j < s; ++j)
      for (int q = 1; q < R; ++q) {
        const long long e = ((long long)j * q) % S;
        const double a = sign * 2.0 * 3.14159265358979323846 * (double)e / (double)S;
        t[(size_t)FP::tw_offset(p) + (size_t)j * FP::tw_pitch(p) + (q - 1)] =
            make_float2((float)std::cos(a), (float)std::sin(a));
      }
  }
  return t;
}

// Base twiddles of passes 1..P-2 for LDS: w_S^j = exp(sign*2*pi*i*j/S),
// j < stride(p), S = span(p), pass after pass.
template <class FP>
constexpr int fixed_lds_tw_offset(int p) {
  int o = 0;
  for (int i = 1; i < p; ++i) o += FP::stride(i);
  return o;
}
template <class FP>
constexpr int fixed_lds_tw_count() {
  return fixed_lds_tw_offset<FP>(FP::kPasses - 1);
}
template <class FP>
std::vector<float2> fixed_lds_twiddles(int sign) {
  std::vector<float2> t((size_t)fixed_lds_tw_count<FP>());
  for (int p = 1; p < FP::kPasses - 1; ++p)
    for (int j = 0; j < FP::stride(p); ++j) {
      const double a = sign * 2.0 * 3.14159265358979323846 * (double)j / (double)FP::span(p);
      t[(size_t)fixed_lds_tw_offset<FP>(p) + j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
  return t;
}

// One in-place DIF pass of a single transform held in LDS by kThreads threads.
// kLdsTw = false: the R-1 twiddles of a butterfly come ready-made from the
// global table twg; true: w^j is read from the LDS base table twl (see
// fixed_lds_twiddles) and its powers are chained.
template <class FP, int p, int kThreads, bool kLdsTw>
__device__ __forceinline__ void fixed_dif_pass(float2 *buf, const float2 *__restrict__ twg,
                                               const float2 *twl, float sgn) {
  constexpr int R = FP::kRadix[p];
  constexpr int S = FP::span(p);
  constexpr int s = FP::stride(p);
  constexpr int nbt = FP::kM / R;
  constexpr int pitch = FP::tw_pitch(p);
  const float2 *__restrict__ tp = twg + FP::tw_offset(p);
  const float2 *tl = twl + fixed_lds_tw_offset<FP>(p);
  // Short strides take the butterflies block-major (consecutive lanes in
  // consecutive blocks, one leg index j): the legs of neighbouring lanes are
  // then a whole block apart, which spreads them over the LDS banks, where
  // j-major runs of s < 32 lanes collide; the shared twiddle read becomes a
  // broadcast.
  constexpr bool kBlockMajor = s < 32;
  constexpr int nblk = FP::kM / S;
  for (int b = threadIdx.x; b < nbt; b += kThreads) {
    const int blk = kBlockMajor ? b % nblk : b / s;
    const int j = kBlockMajor ? b / nblk : b - blk * s;
    float2 *base = buf + blk * S + j;
    float2 w[R - 1];
    if constexpr (kLdsTw) {
      w[0] = tl[j];
#pragma unroll
      for (int q = 1; q < R - 1; ++q) w[q] = cmul(w[q - 1], w[0]);
    } else {
#pragma unroll
      for (int q = 0; q < R - 1; ++q) w[q] = tp[j * pitch + q];
    }
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = base[m * s];
    dft<R>(v, sgn);
#pragma unroll
    for (int q = 1; q < R; ++q) v[q] = cmul(v[q], w[q - 1]);
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * s] = v[q];
  }
  __syncthreads();
}

// Passes p..pEnd-1 with LDS twiddles (by default up to P-2; the last pass
// is left to fixed_last_pass_bin).
template <class FP, int p, int kThreads, int pEnd = FP::kPasses - 1>
__device__ __forceinline__ void fixed_dif_passes_lds(float2 *buf, const float2 *twl, float sgn) {
  if constexpr (p < pEnd) {
    fixed_dif_pass<FP, p, kThreads, true>(buf, nullptr, twl, sgn);
    fixed_dif_passes_lds<FP, p + 1, kThreads, pEnd>(buf, twl, sgn);
  }
}

// Pass P-2 when only the outputs near both ends of the transform are read
// afterwards.  With G = M / (R_{P-2} R_{P-1}) and L = M / R_{P-1}, final bin
// k = m + L q_last and m = lower + G q (lower < G, q the output of this
// pass's butterfly).  Bins k <= n_hi with n_hi < 2G and k >= M - n_neg with
// n_neg <= 2G have m < 2G or m >= L - 2G, so q is 0, 1, R-2 or R-1: those
// four outputs of each odd-R butterfly are formed, twiddled and written, the
// others are neither computed nor stored.  Twiddles: w = w_S^j from the LDS
// base table, w^((R-1)/2) by chaining, w^(R-1) its square, w^(R-2) =
// w^(R-1) conj(w) (|w| = 1).
template <class FP>
constexpr int fixed_edge_group() {
  return FP::kM / (FP::kRadix[FP::kPasses - 2] * FP::kRadix[FP::kPasses - 1]);
}
template <class FP, int kThreads>
__device__ __forceinline__ void fixed_dif_pass_edge(float2 *buf, const float2 *twl, float sgn) {
  constexpr int p = FP::kPasses - 2;
  constexpr int R = FP::kRadix[p];
  static_assert((R & 1) == 1 && R >= 5, "edge pass: odd radix of at least 5");
  constexpr int S = FP::span(p);
  constexpr int s = FP::stride(p);
  constexpr int nbt = FP::kM / R;
  constexpr int H = (R - 1) / 2;
  constexpr RootTable<R> tb{};
  const float2 *tl = twl + fixed_lds_tw_offset<FP>(p);
  constexpr bool kBlockMajor = s < 32;
  constexpr int nblk = FP::kM / S;
  for (int b = threadIdx.x; b < nbt; b += kThreads) {
    const int blk = kBlockMajor ? b % nblk : b / s;
    const int j = kBlockMajor ? b / nblk : b - blk * s;
    float2 *base = buf + blk * S + j;
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = base[m * s];
    float2 t[H + 1], u[H + 1];
    float2 b0 = v[0];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      t[m] = cadd(v[m], v[R - m]);
      u[m] = csub(v[m], v[R - m]);
      b0 = cadd(b0, t[m]);
    }
    float2 out[3];  // b_1, b_{R-2}, b_{R-1}
#pragma unroll
    for (int q = 1; q <= 2; ++q) {
      float2 re = v[0];
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
      if (q == 1) {
        out[0] = cadd(re, iim);
        out[2] = csub(re, iim);
      } else {
        out[1] = csub(re, iim);
      }
    }
    const float2 w = tl[j];
    float2 wh = w;
#pragma unroll
    for (int q = 1; q < H; ++q) wh = cmul(wh, w);
    const float2 wr1 = cmul(wh, wh);
    const float2 wr2 = cmul(wr1, cconj(w));
    base[0] = b0;
    base[s] = cmul(out[0], w);
    base[(R - 2) * s] = cmul(out[1], wr2);
    base[(R - 1) * s] = cmul(out[2], wr1);
  }
  __syncthreads();
}

// Bin k after fixed_fft_head: the direct R-point DFT (R = last radix) of the
// block that holds it, Z[k] = sum_m buf[base + m] w_R^(m t) with
// t = k div (M/R) and base the digit-reversed position of k mod (M/R).
// roots[u] = exp(sgn*2*pi*i*u/R).  kDirect (roots in LDS): every power is
// read as roots[(m t) mod R]; otherwise the powers are chained and re-read
// from roots every 8 terms.
// The block position (base) and the output index t of bin k in the last pass.
template <class FP>
__device__ __forceinline__ void fixed_last_pass_index(int k, int &t, int &base) {
  constexpr int P = FP::kPasses;
  constexpr int R = FP::kRadix[P - 1];
  constexpr int g = FP::kM / R;
  t = k / g;
  int lower = k - t * g;
  base = 0;
#pragma unroll
  for (int p = 0; p < P - 1; ++p) {
    const int q = lower / FP::kRadix[p];
    base += (lower - q * FP::kRadix[p]) * FP::stride(p);
    lower = q;
  }
}

template <class FP, bool kDirect = false>
__device__ __forceinline__ float2 fixed_last_pass_bin(const float2 *buf,
                                                      const float2 *__restrict__ roots, int k);

// fixed_last_pass_bin with the block position and output index given.
template <class FP, bool kDirect = false>
__device__ __forceinline__ float2 fixed_last_pass_at(const float2 *buf,
                                                     const float2 *__restrict__ roots, int t,
                                                     int base) {
  constexpr int R = FP::kRadix[FP::kPasses - 1];
  float2 z = buf[base];
  if constexpr (kDirect) {
    int e = 0;
#pragma unroll
    for (int m = 1; m < R; ++m) {
      e += t;
      if (e >= R) e -= R;
      const float2 v = buf[base + m];
      const float2 wm = roots[e];
      z.x += v.x * wm.x - v.y * wm.y;
      z.y += v.x * wm.y + v.y * wm.x;
    }
  } else {
    const float2 wt = roots[t];
    float2 wm = wt;
    int e = t;
#pragma unroll
    for (int m = 1; m < R; ++m) {
      const float2 v = buf[base + m];
      z.x += v.x * wm.x - v.y * wm.y;
      z.y += v.x * wm.y + v.y * wm.x;
      e += t;
      if (e >= R) e -= R;
      if (m + 1 < R) wm = ((m + 1) & 7) == 0 ? roots[e] : cmul(wm, wt);
    }
  }
  return z;
}

template <class FP, bool kDirect>
__device__ __forceinline__ float2 fixed_last_pass_bin(const float2 *buf,
                                                      const float2 *__restrict__ roots, int k) {
  int t, base;
  fixed_last_pass_index<FP>(k, t, base);
  return fixed_last_pass_at<FP, kDirect>(buf, roots, t, base);
}

// ---------------------------------------------------------------------------
// Whole transforms on a compile-time plan: forward DIF (natural in,
// digit-reversed out) and inverse DIT (digit-reversed in, natural out), for
// kernels that chain several transforms of one row in LDS.  Base twiddles
// w = exp(-2 pi i j / S) (sign -1; the inverse conjugates them): pass 0 reads
// its stride(0) values from a global table (fixed_pass0_twiddles, L2
// resident), passes 1..P-2 from the LDS table of fixed_lds_twiddles(-1); the
// last pass has stride 1 and no twiddles.  The R-1 powers are chained.
// ---------------------------------------------------------------------------
template <class FP>
std::vector<float2> fixed_pass0_twiddles() {
  std::vector<float2> t((size_t)FP::stride(0));
  for (int j = 0; j < FP::stride(0); ++j) {
    const double a = -2.0 * 3.14159265358979323846 * (double)j / (double)FP::kM;
    t[(size_t)j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  return t;
}

// One pass p: DIF (butterfly, then twiddle the outputs) or DIT (twiddle the
// inputs, then butterfly); kInv selects the +1 sign.
template <class FP, int p, int kThreads, bool kDit, bool kInv>
__device__ __forceinline__ void fixed_pass(float2 *buf, const float2 *__restrict__ tw0,
                                           const float2 *twl) {
  constexpr int R = FP::kRadix[p];
  constexpr int S = FP::span(p);
  constexpr int s = FP::stride(p);
  constexpr int nbt = FP::kM / R;
  constexpr bool kBlockMajor = s < 32;  // see fixed_dif_pass
  constexpr int nblk = FP::kM / S;
  constexpr float sgn = kInv ? 1.f : -1.f;
  for (int b = threadIdx.x; b < nbt; b += kThreads) {
    const int blk = kBlockMajor ? b % nblk : b / s;
    const int j = kBlockMajor ? b / nblk : b - blk * s;
    float2 *base = buf + blk * S + j;
    float2 w1 = make_float2(1.f, 0.f);
    if constexpr (s > 1) {
      w1 = p == 0 ? tw0[j] : twl[fixed_lds_tw_offset<FP>(p) + j];
      if (kInv) w1.y = -w1.y;
    }
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = base[m * s];
    // v[q] *= w1^q, the powers chained as they are used (two live values)
    auto twiddle = [&]() {
      float2 wq = w1;
#pragma unroll
      for (int q = 1; q < R; ++q) {
        v[q] = cmul(v[q], wq);
        if (q + 1 < R) wq = cmul(wq, w1);
      }
    };
    if constexpr (s > 1 && kDit) twiddle();
    dft<R>(v, sgn);
    if constexpr (s > 1 && !kDit) twiddle();
#pragma unroll
    for (int q = 0; q < R; ++q) base[q * s] = v[q];
  }
  __syncthreads();
}

template <class FP, int p, int kThreads>
__device__ __forceinline__ void fixed_fft_fwd_from(float2 *buf, const float2 *__restrict__ tw0,
                                                   const float2 *twl) {
  if constexpr (p < FP::kPasses) {
    fixed_pass<FP, p, kThreads, false, false>(buf, tw0, twl);
    fixed_fft_fwd_from<FP, p + 1, kThreads>(buf, tw0, twl);
  }
}
template <class FP, int p, int kThreads>
__device__ __forceinline__ void fixed_fft_inv_from(float2 *buf, const float2 *__restrict__ tw0,
                                                   const float2 *twl) {
  if constexpr (p >= 0) {
    fixed_pass<FP, p, kThreads, true, true>(buf, tw0, twl);
    fixed_fft_inv_from<FP, p - 1, kThreads>(buf, tw0, twl);
  }
}
// Forward (sign -1) DIF: natural order in, digit-reversed (fixed_pos) out.
template <class FP, int kThreads>
__device__ __forceinline__ void fixed_fft_fwd(float2 *buf, const float2 *__restrict__ tw0,
                                              const float2 *twl) {
  fixed_fft_fwd_from<FP, 0, kThreads>(buf, tw0, twl);
}
// Inverse (sign +1, unnormalised) DIT: digit-reversed in, natural out.
template <class FP, int kThreads>
__device__ __forceinline__ void fixed_fft_inv(float2 *buf, const float2 *__restrict__ tw0,
                                              const float2 *twl) {
  fixed_fft_inv_from<FP, FP::kPasses - 1, kThreads>(buf, tw0, twl);
}

// Position of bin k in the digit-reversed result of fixed_fft_fwd.
template <class FP>
__device__ __forceinline__ int fixed_pos(int k) {
  int pos = 0;
#pragma unroll
  for (int p = 0; p < FP::kPasses; ++p) {
    const int q = k / FP::kRadix[p];
    pos += (k - q * FP::kRadix[p]) * FP::stride(p);
    k = q;
  }
  return pos;
}

}  // namespace fxc
