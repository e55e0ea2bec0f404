// sac2spec hot path: per station-window preprocessing, time normalisation,
// spectral whitening and the output spectrum.
//
// Replaces ComputeWorkerBatch (native/sac2spec/worker/compute.cu:8-272) and
// the ~30 element-wise kernels and 4-5 cuFFT passes it launches per batch.
// Rows are station-windows ordered [step][file], file = group*nch + ch; a
// frame (step, group) owns nch consecutive rows that share the normalisation
// and whitening weights (core/normalization.cu:91-213, core/spectrum.cu:178-305).
//
// MI355X design: every per-row stage is one workgroup holding the row's whole
// transform in LDS.  Real data sit in LDS as floats, which is the packed
// complex sequence z[j] = x[2j] + i x[2j+1] of the half-length FFT, so
// packing is free.  Forward transforms are DIF (natural in, digit-reversed
// out), inverse transforms DIT (digit-reversed in, natural out): the
// point-wise filter between them runs in digit-reversed order and no
// permutation pass exists.  Demean/detrend/taper, ×dt, sanitising, |H|^2,
// the normalisation divide and clip, whitening divide and taper, and the
// phase-only option are all folded into those passes.  Sums the reference
// takes in FP64 (mean, trend, running means) stay FP64.
#include "fxc_fft_fixed.hpp"
#include "fxc_fft_large.hpp"
#include "fxc_plan_cache.hpp"
#include "fastxc_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace fxc {

constexpr int kS2sThreads = 1024;
constexpr float kMaxVal = 20.f;        // MAXVAL, include/config.h:13
constexpr float kMinVal = 1e-30f;      // MINVAL as a float argument
constexpr double kMinValD = 1e-30;     // MINVAL in double comparisons

enum { kLoadPlain = 0, kLoadRunabs = 1, kLoadOnebit = 2 };

struct RealTables {
  FftPlanDev plan;    // half-length complex plan (radix structure)
  const float2 *tw_fwd;
  const float2 *tw_inv;
  const float2 *fold;  // exp(-2 pi i k / n), two-level, k <= n/2
};

__device__ __forceinline__ bool cfinite(float2 v) { return isfinite(v.x) && isfinite(v.y); }

// Block-wide double sums (three at once).  Uses a small static LDS area.
__device__ inline void block_sum3(double &a, double &b, double &c) {
  __shared__ double red[3][kS2sThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off);
    b += __shfl_down(b, off);
    c += __shfl_down(c, off);
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
    a += red[0][i];
    b += red[1][i];
    c += red[2][i];
  }
}

// LDS tables: [twiddles fwd | twiddles inv | fold] after `data_slots`.
__device__ inline const float2 *stage_tables(float2 *lds, int data_slots, const RealTables &t,
                                             const float2 **tw_inv, const float2 **fold) {
  const int M = t.plan.n;
  const int ntw = kTwFine + t.plan.ncoarse;
  float2 *a = lds + data_slots;
  float2 *b = a + ntw;
  float2 *c = b + ntw;
  for (int i = threadIdx.x; i < ntw; i += blockDim.x) {
    a[i] = t.tw_fwd[i];
    b[i] = t.tw_inv[i];
  }
  for (int i = threadIdx.x; i < fold_lds_count(M); i += blockDim.x) c[i] = t.fold[i];
  *tw_inv = b;
  *fold = c;
  return a;
}

__host__ __device__ inline int tables_count(int M) { return 2 * tw_lds_count(M) + fold_lds_count(M); }

// After a forward DIF of the packed real sequence (length 2M), visits the
// spectrum pairwise: f(k, X[k], X[M-k]) returns the new values (Y[k],
// Y[M-k]), which are refolded in place for a DIT inverse (w = conj fold).
// k = 0 carries (X[0], X[M]).  Pairs are disjoint, so no barrier is needed
// inside; one barrier at the end.
template <class F>
__device__ inline void spectral_pairs(float2 *lds, const FftPlanDev &plan, const float2 *fold,
                                      const F &f) {
  const int M = plan.n;
  for (int k = threadIdx.x; 2 * k <= M; k += blockDim.x) {
    const int pk = fft_pos(plan, k);
    const int pm = fft_pos(plan, k == 0 ? 0 : M - k);
    const float2 a = lds[pk];
    const float2 b = cconj(lds[pm]);
    float2 xk, xm;
    if (k == 0) {
      xk = make_float2(a.x + a.y, 0.f);  // X[0]
      xm = make_float2(a.x - a.y, 0.f);  // X[M]
    } else {
      const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
      const float2 d = csub(a, b);
      const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
      xk = cadd(e, cmul(fold_lookup(fold, k), o));
      // X[M-k] from the same pair: swap roles of a and b
      const float2 a2 = lds[pm];
      const float2 b2 = cconj(lds[pk]);
      const float2 e2 = make_float2(0.5f * (a2.x + b2.x), 0.5f * (a2.y + b2.y));
      const float2 d2 = csub(a2, b2);
      const float2 o2 = make_float2(0.5f * d2.y, -0.5f * d2.x);
      xm = cadd(e2, cmul(fold_lookup(fold, M - k), o2));
    }
    float2 yk, ym;
    f(k, xk, xm, yk, ym);
    // refold for the inverse: Z[k] = (Y[k] + conj Y[M-k]) + i conj(w_k) (Y[k] - conj Y[M-k])
    if (k == 0) {
      lds[pk] = make_float2(yk.x + ym.x, yk.x - ym.x);
    } else {
      const float2 s1 = cadd(yk, cconj(ym));
      const float2 d1 = cmul(cconj(fold_lookup(fold, k)), csub(yk, cconj(ym)));
      const float2 zk = make_float2(s1.x - d1.y, s1.y + d1.x);
      if (2 * k != M) {
        const float2 s2 = cadd(ym, cconj(yk));
        const float2 d2 = cmul(cconj(fold_lookup(fold, M - k)), csub(ym, cconj(yk)));
        lds[pm] = make_float2(s2.x - d2.y, s2.y + d2.x);
      }
      lds[pk] = zk;
    }
  }
  __syncthreads();
}

// After a forward DIF, emits X[k] (k = 0..M) in natural order: st(k, X[k]).
template <class S>
__device__ inline void spectral_emit(const float2 *lds, const FftPlanDev &plan, const float2 *fold,
                                     const S &st) {
  r2c_unfold(lds, plan, fold, st);
}

// Streaming row accesses (each byte of a row is touched once per kernel):
// nontemporal, so the small tables every workgroup reads stay in L2.
__device__ __forceinline__ float stream_load(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float2 stream_load(const float2 *p) {
  return __builtin_bit_cast(float2, __builtin_nontemporal_load((const unsigned long long *)p));
}
__device__ __forceinline__ void stream_store(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void stream_store(float2 *p, float2 v) {
  __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v), (unsigned long long *)p);
}

// Loads one row of the time series into the float view of LDS applying the
// pending time-domain normalisation (reference order: x/w with 0 for an
// invalid weight, non-finite -> 0, clip to +-MAXVAL; or sign(x)), then zero
// pads to 2M samples.
__device__ inline void load_row(float *ldsf, int twoM, const float *__restrict__ x, int seg,
                                int mode, const float *__restrict__ w) {
  constexpr int kB = 8;  // samples whose loads are issued before the first is used
  for (int i0 = threadIdx.x; i0 < twoM; i0 += kB * blockDim.x) {
    float xv[kB], wv[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int i = i0 + u * blockDim.x;
      xv[u] = 0.f;
      wv[u] = 0.f;
      if (i < seg && i < twoM) {
        xv[u] = stream_load(x + i);
        if (mode == kLoadRunabs) wv[u] = w[i];
      }
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int i = i0 + u * blockDim.x;
      if (i >= twoM) continue;
      float v = 0.f;
      if (i < seg) {
        v = xv[u];
        if (mode == kLoadRunabs) {
          const float d = wv[u];
          v = (!isfinite(d) || d < kMinValD) ? 0.f : v / d;
          if (!isfinite(v)) v = 0.f;
          v = v > kMaxVal ? kMaxVal : (v < -kMaxVal ? -kMaxVal : v);
        } else if (mode == kLoadOnebit) {
          v = v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f);
        }
      }
      ldsf[i] = v;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// K1: window cut + preprocess + broad-band prefilter (compute.cu:34-86).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kS2sThreads)
s2s_prefilter_kernel(const float *__restrict__ traces, int npts, int nfiles,
                     const int *__restrict__ steps, int shift, int seg, int taper_size,
                     RealTables tf, const float *__restrict__ resp0, float dt, double inv_weight,
                     float *__restrict__ xout) {
  extern __shared__ float2 lds[];
  float *ldsf = (float *)lds;
  const FftPlanDev &plan = tf.plan;
  const int M = plan.n;
  const float2 *tw_inv, *fold;
  const float2 *tw_fwd = stage_tables(lds, M + 1, tf, &tw_inv, &fold);
  const int row = blockIdx.x;
  const int step = row / nfiles, file = row - step * nfiles;
  const long long off = (long long)steps[step] * shift;
  const float *src = traces + (size_t)file * npts;
  // load + non-finite -> 0 + FP64 sum; a batch of samples' loads is
  // issued before the first is used
  double s = 0.0, unused1 = 0.0, unused2 = 0.0;
  constexpr int kB = 8;
  for (int i0 = threadIdx.x; i0 < 2 * M; i0 += kB * blockDim.x) {
    float xv[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int i = i0 + u * blockDim.x;
      xv[u] = (i < seg && off + i < npts) ? stream_load(src + off + i) : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int i = i0 + u * blockDim.x;
      if (i >= 2 * M) continue;
      const float v = isfinite(xv[u]) ? xv[u] : 0.f;
      ldsf[i] = v;
      if (i < seg) s += v;
    }
  }
  block_sum3(s, unused1, unused2);
  // demean: x -= sum/W (double), then sums for the linear trend
  const double mean = s / seg;
  double s1 = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    const float v = (float)((double)ldsf[i] - mean);
    ldsf[i] = v;
    s1 += v;
    si += (double)v * i;
  }
  unused1 = 0.0;
  block_sum3(s1, si, unused1);
  const double W = seg;
  const double a12 = 0.5 * W * (W - 1);
  const double a11 = a12 * (2 * W - 1) / 3.;
  const double a22 = W;
  const double det = a11 * a22 - a12 * a12;
  const double ca = (a22 * si - a12 * s1) / det;
  const double cb = (a11 * s1 - a12 * si) / det;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    float v = (float)((double)ldsf[i] - (ca * i + cb));
    float factor = 1.f;
    if (i < taper_size)
      factor = (float)(0.5 * (1.0 - cos(M_PI * i / taper_size)));
    else if (i >= seg - taper_size)
      factor = (float)(0.5 * (1.0 - cos(M_PI * (seg - i) / taper_size)));
    ldsf[i] = v * factor;
  }
  __syncthreads();
  lds_fft(lds, 1, plan, tw_fwd, -1.f);
  spectral_pairs(lds, plan, fold, [&](int k, float2 xk, float2 xm, float2 &yk, float2 &ym) {
    xk = cscale(xk, dt);
    xm = cscale(xm, dt);
    if (!cfinite(xk)) xk = make_float2(0.f, 0.f);
    if (!cfinite(xm)) xm = make_float2(0.f, 0.f);
    yk = cscale(xk, resp0[k]);
    ym = cscale(xm, resp0[k == 0 ? M : M - k]);
  });
  lds_fft_dit(lds, 1, plan, tw_inv, 1.f);
  float *y = xout + (size_t)row * seg;
  for (int i = threadIdx.x; i < seg; i += blockDim.x)
    stream_store(y + i, (float)((double)ldsf[i] * inv_weight));
}

// ---------------------------------------------------------------------------
// Running-mean weights shared by a frame's channels:
//   w[t] = sum_ch (1/W) sum_{i=t-W/2}^{t-W/2+W-1} |x_ch[i]|  (clipped edges)
// (smoothRowsRollingKernel, kernels/real_matrix.cu:155-191, summed by
// sum2DKernel and clamped by clampmin2DKernel).  Each thread slides an FP64
// window over a contiguous block of outputs.  `cplx` selects |X| of complex
// rows (whitening) instead of |x|.
// ---------------------------------------------------------------------------
constexpr int kWeightChunk = 128;  // consecutive outputs per thread

template <bool kCplx>
__global__ void __launch_bounds__(256)
s2s_weight_kernel(const void *__restrict__ data, long long row_pitch, int width, int nch, int win,
                  float *__restrict__ wout, long long w_pitch) {
  const int frame = blockIdx.y;
  const int t0 = (blockIdx.x * blockDim.x + threadIdx.x) * kWeightChunk;
  if (t0 >= width) return;
  const int t1 = min(width, t0 + kWeightChunk);
  const double inv = 1.0 / (double)win;
  float *w = wout + (long long)frame * w_pitch;
  for (int ch = 0; ch < nch; ++ch) {
    const long long r = (long long)frame * nch + ch;
    auto val = [&](int i) -> double {
      if (kCplx) {
        const float2 z = ((const float2 *)data)[r * row_pitch + i];
        return (double)sqrtf(z.x * z.x + z.y * z.y);
      }
      return (double)fabsf(((const float *)data)[r * row_pitch + i]);
    };
    int start = t0 - win / 2;
    int end = start + win - 1;
    double sum = 0.0;
    for (int i = max(start, 0); i <= min(end, width - 1); ++i) sum += val(i);
    for (int t = t0; t < t1; ++t) {
      const float v = (float)(inv * sum);
      w[t] = ch == 0 ? v : w[t] + v;  // this thread owns outputs t0..t1-1
      if (start >= 0 && start < width) sum -= val(start);
      ++start;
      ++end;
      if (end >= 0 && end < width) sum += val(end);
    }
  }
  for (int t = t0; t < t1; ++t) {
    float v = w[t];
    if (!isfinite(v) || v < kMinVal) v = kMinVal;
    w[t] = v;
  }
}

// Tiled form of the same weights.  A workgroup owns kWTile consecutive
// outputs of one frame; per channel it stages v[j] = |x[t0 - win/2 + j]|
// (0 outside the row), j < kWTile + win - 1, in LDS as doubles (coalesced
// loads), forms their inclusive prefix S in FP64 and takes every window sum
// as S[k + win - 1] - S[k - 1].  The prefix is local to the tile, so the
// difference loses at most ~1e-14 relative -- far below the float rounding
// of the result -- where the rolling form of s2s_weight_kernel re-reads
// each thread's window with scattered loads.  Channels are added in float
// in channel order, as sum2DKernel adds the smoothed rows.
constexpr int kWTile = 4096;
constexpr int kWThreads = 256;
constexpr int kWPerThread = kWTile / kWThreads;
constexpr int kWTileMaxWin = 12000;  // LDS: (kWTile + win - 1) doubles

template <bool kCplx>
__global__ void __launch_bounds__(kWThreads)
s2s_weight_tile_kernel(const void *__restrict__ data, long long row_pitch, int width, int nch,
                       int win, float *__restrict__ wout, long long w_pitch) {
  extern __shared__ double S[];
  __shared__ double wave_tot[kWThreads / 64];
  const long long frame = blockIdx.x;
  const int t0 = blockIdx.y * kWTile;
  const int T = min(kWTile, width - t0);
  const int L = T + win - 1;
  const int base = t0 - win / 2;
  const double inv = 1.0 / (double)win;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int per = (L + kWThreads - 1) / kWThreads;
  const int a = min(L, tid * per), b = min(L, a + per);
  float acc[kWPerThread];
#pragma unroll
  for (int m = 0; m < kWPerThread; ++m) acc[m] = 0.f;
  for (int ch = 0; ch < nch; ++ch) {
    const long long r = frame * nch + ch;
    __syncthreads();  // the previous channel's reads of S are done
    for (int j = tid; j < L; j += kWThreads) {
      const int i = base + j;
      double v = 0.0;
      if (i >= 0 && i < width) {
        if (kCplx) {
          const float2 z = ((const float2 *)data)[r * row_pitch + i];
          v = (double)sqrtf(z.x * z.x + z.y * z.y);
        } else {
          v = (double)fabsf(((const float *)data)[r * row_pitch + i]);
        }
      }
      S[j] = v;
    }
    __syncthreads();
    double run = 0.0;
    for (int j = a; j < b; ++j) run += S[j];
    double incl = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    if (lane == 63) wave_tot[wid] = incl;
    __syncthreads();
    double pre = incl - run;
    for (int w = 0; w < wid; ++w) pre += wave_tot[w];
    for (int j = a; j < b; ++j) {
      pre += S[j];
      S[j] = pre;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kWPerThread; ++m) {
      const int k = tid + m * kWThreads;
      if (k < T) {
        const double sum = S[k + win - 1] - (k > 0 ? S[k - 1] : 0.0);
        acc[m] += (float)(inv * sum);
      }
    }
  }
  float *w = wout + frame * w_pitch + t0;
#pragma unroll
  for (int m = 0; m < kWPerThread; ++m) {
    const int k = tid + m * kWThreads;
    if (k < T) {
      float v = acc[m];
      if (!isfinite(v) || v < kMinVal) v = kMinVal;
      w[k] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Whitening forward: load (normalised) row, R2C(seg) x dt, sanitise; write
// X[k] and |X[k]|, k < seg/2+1 (core/spectrum.cu:178-230 up to the weight).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kS2sThreads)
s2s_whiten_fwd_kernel(const float *__restrict__ x, int seg, int nch, int mode,
                      const float *__restrict__ w, RealTables ts, float dt,
                      float2 *__restrict__ X, float *__restrict__ amp, long long fpitch) {
  extern __shared__ float2 lds[];
  const FftPlanDev &plan = ts.plan;
  const int M = plan.n;
  const float2 *tw_inv, *fold;
  const float2 *tw_fwd = stage_tables(lds, M + 1, ts, &tw_inv, &fold);
  const int row = blockIdx.x;
  const int frame = row / nch;
  load_row((float *)lds, 2 * M, x + (size_t)row * seg, seg, mode,
           w ? w + (size_t)frame * seg : nullptr);
  lds_fft(lds, 1, plan, tw_fwd, -1.f);
  float2 *Xr = X + (size_t)row * fpitch;
  float *ar = amp + (size_t)row * fpitch;
  spectral_emit(lds, plan, fold, [&](int k, float2 v) {
    v = cscale(v, dt);
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
    stream_store(Xr + k, v);
    stream_store(ar + k, sqrtf(v.x * v.x + v.y * v.y));
  });
}

// Spectral taper factor (specTaper2DKernel, kernels/taper.cu:8-69, np = 1).
__device__ inline float spec_taper(int col, int i1, int i2, int i3, int i4) {
  if (col < i1) return 0.f;
  if (col < i2) return (float)((1.0 - cos(M_PI / (i2 - i1) * (col - i1))) / 2.0);
  if (col >= i3 && col < i4) return (float)((1.0 + cos(M_PI / (i4 - i3) * (col - i3))) / 2.0);
  if (col >= i4) return 0.f;
  return 1.f;
}

// ---------------------------------------------------------------------------
// Whitening inverse (+ optionally the output spectrum): X/ww (0 if the weight
// is invalid), taper, sanitise, C2R(seg) x 1/(seg dt) (spectrum.cu:231-305,
// compute.cu:203-217); then either store x, or zero-pad to nfft_out, R2C x dt,
// sanitise, optional phase-only and store nspec bins (compute.cu:226-264).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kS2sThreads)
s2s_whiten_inv_kernel(const float2 *__restrict__ X, long long fpitch, const float *__restrict__ ww,
                      int nch, int seg, int i1, int i2, int i3, int i4, RealTables ts,
                      double inv_weight_seg, float *__restrict__ xout, RealTables to, float dt,
                      int phase_only, float2 *__restrict__ out, int nspec) {
  extern __shared__ float2 lds[];
  const FftPlanDev &ps = ts.plan;
  const int Ms = ps.n;
  const bool do_out = out != nullptr;
  const int Mo = do_out ? to.plan.n : Ms;
  const int data = (Ms > Mo ? Ms : Mo) + 1;
  const float2 *tws_inv, *fold_s;
  (void)stage_tables(lds, data, ts, &tws_inv, &fold_s);
  const float2 *two_inv = nullptr, *fold_o = nullptr, *two_fwd = nullptr;
  if (do_out) two_fwd = stage_tables(lds, data + tables_count(Ms), to, &two_inv, &fold_o);
  __syncthreads();
  const int row = blockIdx.x;
  const int frame = row / nch;
  const float2 *Xr = X + (size_t)row * fpitch;
  const float *wr = ww + (size_t)frame * fpitch;
  const int nbin = Ms + 1;
  const int t4 = i4 < nbin ? i4 : nbin;
  auto whitened_from = [&](int k, float2 v, float d) -> float2 {
    v = (!isfinite(d) || d < kMinValD) ? make_float2(0.f, 0.f) : make_float2(v.x / d, v.y / d);
    v = cscale(v, spec_taper(k, i1, i2, i3, t4));
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
    return v;
  };
  // fold into digit-reversed positions for the DIT inverse; a batch of
  // bins' loads is issued before the first is used
  constexpr int kB = 4;
  for (int k0 = threadIdx.x; k0 < Ms; k0 += kB * blockDim.x) {
    float2 xa[kB], xb[kB];
    float da[kB], db[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 + u * blockDim.x;
      if (k < Ms) {
        xa[u] = stream_load(Xr + k);
        da[u] = wr[k];
        xb[u] = stream_load(Xr + Ms - k);
        db[u] = wr[Ms - k];
      }
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = k0 + u * blockDim.x;
      if (k >= Ms) continue;
      float2 a = whitened_from(k, xa[u], da[u]);
      float2 b = cconj(whitened_from(Ms - k, xb[u], db[u]));
      if (k == 0) {
        a.y = 0.f;
        b.y = 0.f;
      }
      const float2 s = cadd(a, b);
      const float2 d = cmul(cconj(fold_lookup(fold_s, k)), csub(a, b));
      lds[fft_pos(ps, k)] = make_float2(s.x - d.y, s.y + d.x);
    }
  }
  __syncthreads();
  lds_fft_dit(lds, 1, ps, tws_inv, 1.f);
  float *ldsf = (float *)lds;
  if (!do_out) {
    float *y = xout + (size_t)row * seg;
    for (int i = threadIdx.x; i < seg; i += blockDim.x)
      stream_store(y + i, (float)((double)ldsf[i] * inv_weight_seg));
    return;
  }
  for (int i = threadIdx.x; i < 2 * Mo; i += blockDim.x)
    ldsf[i] = i < seg ? (float)((double)ldsf[i] * inv_weight_seg) : 0.f;
  __syncthreads();
  lds_fft(lds, 1, to.plan, two_fwd, -1.f);
  float2 *o = out + (size_t)row * nspec;
  spectral_emit(lds, to.plan, fold_o, [&](int k, float2 v) {
    if (k >= nspec) return;
    v = cscale(v, dt);
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
    if (phase_only) {
      const float a = hypotf(v.x, v.y);
      v = (!isfinite(a) || a <= kMinVal) ? make_float2(0.f, 0.f) : cscale(v, 1.f / a);
      if (!cfinite(v)) v = make_float2(0.f, 0.f);
    }
    stream_store(o + k, v);
  });
}

// Output spectrum without whitening-after: load (normalised) row, zero-pad,
// R2C(nfft_out) x dt, sanitise, optional phase-only (compute.cu:226-264).
__global__ void __launch_bounds__(kS2sThreads)
s2s_output_kernel(const float *__restrict__ x, int seg, int nch, int mode,
                  const float *__restrict__ w, RealTables to, float dt, int phase_only,
                  float2 *__restrict__ out, int nspec) {
  extern __shared__ float2 lds[];
  const FftPlanDev &plan = to.plan;
  const int M = plan.n;
  const float2 *tw_inv, *fold;
  const float2 *tw_fwd = stage_tables(lds, M + 1, to, &tw_inv, &fold);
  const int row = blockIdx.x;
  const int frame = row / nch;
  load_row((float *)lds, 2 * M, x + (size_t)row * seg, seg, mode,
           w ? w + (size_t)frame * seg : nullptr);
  lds_fft(lds, 1, plan, tw_fwd, -1.f);
  float2 *o = out + (size_t)row * nspec;
  spectral_emit(lds, plan, fold, [&](int k, float2 v) {
    if (k >= nspec) return;
    v = cscale(v, dt);
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
    if (phase_only) {
      const float a = hypotf(v.x, v.y);
      v = (!isfinite(a) || a <= kMinVal) ? make_float2(0.f, 0.f) : cscale(v, 1.f / a);
      if (!cfinite(v)) v = make_float2(0.f, 0.f);
    }
    stream_store(o + k, v);
  });
}

// RUN-ABS-MF base spectrum: R2C(nfft_filter) x dt, sanitised (normalization.cu:285-329).
__global__ void __launch_bounds__(kS2sThreads)
s2s_mf_base_kernel(const float *__restrict__ x, int seg, RealTables tf, float dt,
                   float2 *__restrict__ base, long long bpitch) {
  extern __shared__ float2 lds[];
  const FftPlanDev &plan = tf.plan;
  const int M = plan.n;
  const float2 *tw_inv, *fold;
  const float2 *tw_fwd = stage_tables(lds, M + 1, tf, &tw_inv, &fold);
  const int row = blockIdx.x;
  load_row((float *)lds, 2 * M, x + (size_t)row * seg, seg, kLoadPlain, nullptr);
  lds_fft(lds, 1, plan, tw_fwd, -1.f);
  float2 *b = base + (size_t)row * bpitch;
  spectral_emit(lds, plan, fold, [&](int k, float2 v) {
    v = cscale(v, dt);
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
    b[k] = v;
  });
}

// RUN-ABS-MF band i: C2R(base x |H_i|^2) x 1/(nfft_filter dt), first seg
// samples (normalization.cu:331-370).
__global__ void __launch_bounds__(kS2sThreads)
s2s_mf_band_kernel(const float2 *__restrict__ base, long long bpitch, RealTables tf,
                   const float *__restrict__ resp, double inv_weight, int seg,
                   float *__restrict__ xout) {
  extern __shared__ float2 lds[];
  const FftPlanDev &plan = tf.plan;
  const int M = plan.n;
  const float2 *tw_inv, *fold;
  (void)stage_tables(lds, M + 1, tf, &tw_inv, &fold);
  __syncthreads();
  const int row = blockIdx.x;
  const float2 *b = base + (size_t)row * bpitch;
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    float2 a = cscale(b[k], resp[k]);
    float2 c = cconj(cscale(b[M - k], resp[M - k]));
    if (k == 0) {
      a.y = 0.f;
      c.y = 0.f;
    }
    const float2 s = cadd(a, c);
    const float2 d = cmul(cconj(fold_lookup(fold, k)), csub(a, c));
    lds[fft_pos(plan, k)] = make_float2(s.x - d.y, s.y + d.x);
  }
  __syncthreads();
  lds_fft_dit(lds, 1, plan, tw_inv, 1.f);
  const float *ldsf = (const float *)lds;
  float *y = xout + (size_t)row * seg;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) y[i] = (float)((double)ldsf[i] * inv_weight);
}

// total += runabs-normalised band signal (normalization.cu:371-385).
// Workgroup b covers samples [(b mod nbr) blockDim, +blockDim) of row
// b / nbr, so the row comes from one division per workgroup rather than
// one per sample.
__global__ void s2s_mf_accum_kernel(const float *__restrict__ xb, const float *__restrict__ w,
                                    int seg, int nch, int nbr, float *__restrict__ total) {
  const int row = (int)(blockIdx.x / (unsigned)nbr);
  const int t = (int)(blockIdx.x - (unsigned)row * nbr) * blockDim.x + threadIdx.x;
  if (t >= seg) return;
  const long long i = (long long)row * seg + t;
  const long long frame = row / nch;
  float v = xb[i];
  const float d = w[frame * seg + t];
  v = (!isfinite(d) || d < kMinValD) ? 0.f : v / d;
  if (!isfinite(v)) v = 0.f;
  v = v > kMaxVal ? kMaxVal : (v < -kMaxVal ? -kMaxVal : v);
  total[i] = total[i] + v;
}

// ---------------------------------------------------------------------------
// Long windows (transforms beyond one workgroup's LDS, e.g. dt 0.05 s: seg
// 72000, filter FFT 72576, output FFT 74088).  The same steps in the same
// order as the fused kernels above, as element-wise passes over HBM rows
// around the four-step transforms of fxc_fft_large.
// ---------------------------------------------------------------------------
// Window cut, non-finite -> 0, demean, detrend (FP64 sums) and edge taper of
// one row into x[row][seg] (the front of s2s_prefilter_kernel).
__global__ void __launch_bounds__(kS2sThreads)
s2sL_pre_kernel(const float *__restrict__ traces, int npts, int nfiles,
                const int *__restrict__ steps, int shift, int seg, int taper_size,
                float *__restrict__ xout) {
  const int row = blockIdx.x;
  const int step = row / nfiles, file = row - step * nfiles;
  const long long off = (long long)steps[step] * shift;
  const float *src = traces + (size_t)file * npts;
  float *y = xout + (size_t)row * seg;
  double s = 0.0, unused1 = 0.0, unused2 = 0.0;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    float v = off + i < npts ? src[off + i] : 0.f;
    if (!isfinite(v)) v = 0.f;
    y[i] = v;
    s += v;
  }
  block_sum3(s, unused1, unused2);
  const double mean = s / seg;
  double s1 = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    const float v = (float)((double)y[i] - mean);
    y[i] = v;
    s1 += v;
    si += (double)v * i;
  }
  unused1 = 0.0;
  block_sum3(s1, si, unused1);
  const double W = seg;
  const double a12 = 0.5 * W * (W - 1);
  const double a11 = a12 * (2 * W - 1) / 3.;
  const double a22 = W;
  const double det = a11 * a22 - a12 * a12;
  const double ca = (a22 * si - a12 * s1) / det;
  const double cb = (a11 * s1 - a12 * si) / det;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    const float v = (float)((double)y[i] - (ca * i + cb));
    float factor = 1.f;
    if (i < taper_size)
      factor = (float)(0.5 * (1.0 - cos(M_PI * i / taper_size)));
    else if (i >= seg - taper_size)
      factor = (float)(0.5 * (1.0 - cos(M_PI * (seg - i) / taper_size)));
    y[i] = v * factor;
  }
}

// X[k] = sanitise(X[k] dt) (times resp[k] when given; |X| into amp when
// given), k < nbin, every row.
__global__ void s2sL_spec_kernel(float2 *__restrict__ X, long long pitch, int nbin, float dt,
                                 const float *__restrict__ resp, float *__restrict__ amp,
                                 long long apitch) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbin) return;
  const long long row = blockIdx.y;
  float2 v = cscale(X[row * pitch + k], dt);
  if (!cfinite(v)) v = make_float2(0.f, 0.f);
  if (resp) v = cscale(v, resp[k]);
  X[row * pitch + k] = v;
  if (amp) amp[row * apitch + k] = sqrtf(v.x * v.x + v.y * v.y);
}

// Pending time-domain normalisation applied in place (load_row's rules).
__global__ void s2sL_norm_kernel(float *__restrict__ x, int seg, int nch, int mode,
                                 const float *__restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= seg) return;
  const long long row = blockIdx.y;
  float v = x[row * seg + i];
  if (mode == kLoadRunabs) {
    const float d = w[(row / nch) * seg + i];
    v = (!isfinite(d) || d < kMinValD) ? 0.f : v / d;
    if (!isfinite(v)) v = 0.f;
    v = v > kMaxVal ? kMaxVal : (v < -kMaxVal ? -kMaxVal : v);
  } else if (mode == kLoadOnebit) {
    v = v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f);
  }
  x[row * seg + i] = v;
}

// Whitening: X / ww (0 for an invalid weight), spectral taper, sanitise.
__global__ void s2sL_whiten_kernel(float2 *__restrict__ X, long long fpitch, int nbin,
                                   const float *__restrict__ ww, int nch, int i1, int i2, int i3,
                                   int i4) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbin) return;
  const long long row = blockIdx.y;
  float2 v = X[row * fpitch + k];
  const float d = ww[(row / nch) * fpitch + k];
  v = (!isfinite(d) || d < kMinValD) ? make_float2(0.f, 0.f) : make_float2(v.x / d, v.y / d);
  v = cscale(v, spec_taper(k, i1, i2, i3, i4 < nbin ? i4 : nbin));
  if (!cfinite(v)) v = make_float2(0.f, 0.f);
  X[row * fpitch + k] = v;
}

// One band of RUN-ABS-MF: dst[k] = base[k] resp[k].
__global__ void s2sL_band_kernel(const float2 *__restrict__ base, long long bpitch, int nbin,
                                 const float *__restrict__ resp, float2 *__restrict__ dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbin) return;
  const long long row = blockIdx.y;
  dst[row * bpitch + k] = cscale(base[row * bpitch + k], resp[k]);
}

// Output spectrum: x dt, sanitise, optional phase-only, in place.
__global__ void s2sL_out_kernel(float2 *__restrict__ out, int nspec, float dt, int phase_only) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nspec) return;
  const long long row = blockIdx.y;
  float2 v = cscale(out[row * nspec + k], dt);
  if (!cfinite(v)) v = make_float2(0.f, 0.f);
  if (phase_only) {
    const float a = hypotf(v.x, v.y);
    v = (!isfinite(a) || a <= kMinVal) ? make_float2(0.f, 0.f) : cscale(v, 1.f / a);
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
  }
  out[row * nspec + k] = v;
}

// ---------------------------------------------------------------------------
// Fused one-channel path at the production lengths (3600 s windows at 0.1 s:
// seg 36000, prefilter FFT 36288, output FFT 37044).  The whole chain of one
// station-window runs in one workgroup's LDS on compile-time plans, in at
// most three kernels instead of five, and only the trace window, the
// whitening spectrum X (when whitening) and the output spectrum touch HBM:
//   front: window cut, preprocess, prefilter, [normalise], R2C(seg) -> X  or
//          R2C(nfft_out) -> output
//   back:  X -> whitening weights (running mean of |X| over the bins), X/w,
//          taper, C2R(seg), [normalise], R2C(nfft_out) -> output or
//          R2C(seg) -> X (whitening before and after)
// A frame of one channel owns its weights, so the running means over time
// (RUN-ABS) and over bins (whitening) are taken in LDS: each thread slides
// an FP64 window over a contiguous block of outputs, as the reference's
// smoothRowsRollingKernel (kernels/real_matrix.cu:155-191) with its clamp.
// ---------------------------------------------------------------------------
// Radix orders keep the passes whose legs sit within one block (stride < 32,
// taken block-major) on odd spans (81, 9; 75, 5; 63, 9): neighbouring lanes
// then address blocks an odd number of words apart, which spreads them over
// the LDS banks, where even spans such as 108 put 8 to 16 lanes on one bank.
// Four passes measured faster than five better-balanced ones (6, 6, 8, 9, 7
// and 18, 8, 5, 5, 5: 15.2 against 14.8 ms for 1000 x 24 station-windows).
using S2sFixF = FixedPlan<18144, 16, 14, 9, 9>;   // prefilter: nfft 36288
using S2sFixS = FixedPlan<18000, 16, 15, 15, 5>;  // window: seg 36000
using S2sFixO = FixedPlan<18522, 6, 7, 7, 7, 9>;   // output: nfft 37044
static_assert(S2sFixF::valid() && S2sFixS::valid() && S2sFixO::valid(), "plans");
constexpr int kFusedPer = 37;   // RUN-ABS outputs per thread (odd: conflict-free LDS reads)
constexpr int kFusedPerB = 19;  // whitening bins per thread
static_assert(kFusedPer * kS2sThreads >= 2 * S2sFixS::kM, "RUN-ABS blocks cover seg");
static_assert(kFusedPerB * kS2sThreads >= S2sFixS::kM + 1, "whitening blocks cover the bins");
// float2 slots after the tables: the block sums of the running means
// (kS2sThreads floats) and, in the back kernel's whitening, the two lane
// exchanges of whiten_mean_lanes (2 x 16 x 2 x kFusedPerB floats)
constexpr int kFusedBsSlots =
    kS2sThreads / 2 > 2 * 16 * kFusedPerB ? kS2sThreads / 2 : 2 * 16 * kFusedPerB;
static_assert(kFusedPerB * kS2sThreads >= S2sFixS::kM + 1, "whitening blocks cover seg/2+1");

struct FixedTables {
  const float2 *tw0;   // pass-0 base twiddles (global)
  const float2 *twl;   // passes 1..P-2 base twiddles (copied to LDS)
  const float2 *fold;  // exp(-2 pi i k / 2M), k <= M, two-level (copied to LDS)
};

template <class FP>
__device__ __forceinline__ void stage_fixed(float2 *twl_lds, float2 *fold_lds, const FixedTables &t) {
  for (int i = threadIdx.x; i < fixed_lds_tw_count<FP>(); i += blockDim.x) twl_lds[i] = t.twl[i];
  for (int i = threadIdx.x; i < fold_lds_count(FP::kM); i += blockDim.x) fold_lds[i] = t.fold[i];
}

// fold_lds_count as a constant expression
constexpr int fused_fold_count(int M) { return kTwFine + (M + 1 + kTwFine - 1) / kTwFine; }

// [data][LDS twiddles A][LDS twiddles B][fold A][fold B][block sums: kS2sThreads floats]
template <class FPA, class FPB>
constexpr int fused_lds_count() {
  return (FPA::kM > FPB::kM ? FPA::kM : FPB::kM) + 1 + fixed_lds_tw_count<FPA>() +
         fixed_lds_tw_count<FPB>() + fused_fold_count(FPA::kM) + fused_fold_count(FPB::kM) +
         kFusedBsSlots;
}

// After fixed_fft_fwd of the packed real row: X[k], k = 0..M, to st(k, X[k])
// (the unfold of r2c_unfold at the plan's digit-reversed positions).
template <class FP, class St>
__device__ __forceinline__ void fixed_r2c_emit(const float2 *lds, const float2 *fold, const St &st,
                                               int sh = 0) {
  constexpr int M = FP::kM;
  // bins start at -sh: with sh the destination row's offset within a
  // 128-byte line (in float2), 16 consecutive lanes store one whole line
  for (int k = (int)threadIdx.x - sh; k <= M; k += blockDim.x) {
    if (k < 0) continue;
    const float2 a = lds[fixed_pos<FP>(k == M ? 0 : k)];
    const float2 b = cconj(lds[fixed_pos<FP>(k == 0 ? 0 : M - k)]);
    const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
    const float2 d = csub(a, b);
    const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
    float2 x = cadd(e, cmul(fold_lookup(fold, k), o));
    if (k == 0 || k == M) x.y = 0.f;
    st(k, x);
  }
}

// Pair trips of fixed_spectral_pairs per thread: pairs k = 0..M/2 over
// kS2sThreads threads.
template <class FP>
constexpr int fixed_pair_trips() {
  return (FP::kM / 2 + kS2sThreads) / kS2sThreads;
}

// spectral_pairs at the plan's digit-reversed positions: f(it, k, X[k],
// X[M-k]) -> (Y[k], Y[M-k]) for the it-th pair of the thread, k = tid + it
// kS2sThreads, refolded in place for fixed_fft_inv.
template <class FP, class F>
__device__ __forceinline__ void fixed_spectral_pairs(float2 *lds, const float2 *fold, const F &f) {
  constexpr int M = FP::kM;
#pragma unroll
  for (int it = 0; it < fixed_pair_trips<FP>(); ++it) {
    const int k = (int)threadIdx.x + it * kS2sThreads;
    if (2 * k > M) continue;
    const int pk = fixed_pos<FP>(k);
    const int pm = fixed_pos<FP>(k == 0 ? 0 : M - k);
    const float2 a = lds[pk];
    const float2 b = cconj(lds[pm]);
    float2 xk, xm;
    if (k == 0) {
      xk = make_float2(a.x + a.y, 0.f);
      xm = make_float2(a.x - a.y, 0.f);
    } else {
      const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
      const float2 d = csub(a, b);
      const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
      xk = cadd(e, cmul(fold_lookup(fold, k), o));
      const float2 a2 = lds[pm];
      const float2 b2 = cconj(lds[pk]);
      const float2 e2 = make_float2(0.5f * (a2.x + b2.x), 0.5f * (a2.y + b2.y));
      const float2 d2 = csub(a2, b2);
      const float2 o2 = make_float2(0.5f * d2.y, -0.5f * d2.x);
      xm = cadd(e2, cmul(fold_lookup(fold, M - k), o2));
    }
    float2 yk, ym;
    f(it, k, xk, xm, yk, ym);
    if (k == 0) {
      lds[pk] = make_float2(yk.x + ym.x, yk.x - ym.x);
    } else {
      const float2 s1 = cadd(yk, cconj(ym));
      const float2 d1 = cmul(cconj(fold_lookup(fold, k)), csub(yk, cconj(ym)));
      const float2 zk = make_float2(s1.x - d1.y, s1.y + d1.x);
      if (2 * k != M) {
        const float2 s2 = cadd(ym, cconj(yk));
        const float2 d2 = cmul(cconj(fold_lookup(fold, M - k)), csub(ym, cconj(yk)));
        lds[pm] = make_float2(s2.x - d2.y, s2.y + d2.x);
      }
      lds[pk] = zk;
    }
  }
  __syncthreads();
}

// Window-sum accumulators of the running means.  Acc64: an FP64 sum, the
// reference's arithmetic (kernels/real_matrix.cu:155-191).  AccComp (spec
// mode, fxc_s2s_set_sums): a float sum with Neumaier compensation, so the
// running means use no FP64; the window mean is (s + c) * (1/win) in float.
struct Acc64 {
  double s = 0.0;
  __device__ __forceinline__ void add(float x) { s += (double)x; }
  __device__ __forceinline__ void sub(float x) { s -= (double)x; }
  __device__ __forceinline__ float mean(int win) const { return (float)((1.0 / (double)win) * s); }
  __device__ __forceinline__ float total() const { return (float)s; }
};
struct AccComp {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float t = s + x;
    c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void sub(float x) { add(-x); }
  __device__ __forceinline__ float mean(int win) const { return (s + c) * (1.f / (float)win); }
  __device__ __forceinline__ float total() const { return s + c; }
};

// Running mean w_t = (1/win) sum_{i = t - win/2}^{t - win/2 + win - 1} a(i)
// over i in [0, n) (window sum in Acc, float result, clamped to >= MINVAL),
// for the kPer contiguous outputs t0 = threadIdx.x * kPer + m of this
// thread, each handed to use(m, t, w_t).
template <int kPer, class Acc, class A, class U>
__device__ __forceinline__ void sliding_mean(int n, int win, const A &a, const U &use) {
  const int t0 = (int)threadIdx.x * kPer;
  if (t0 >= n) return;
  int start = t0 - win / 2, end = start + win - 1;
  Acc sum;
  for (int i = max(start, 0); i <= min(end, n - 1); ++i) sum.add(a(i));
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    const int t = t0 + m;
    if (t < n) {
      float w = sum.mean(win);
      if (!isfinite(w) || w < kMinVal) w = kMinVal;
      use(m, t, w);
      if (start >= 0 && start < n) sum.sub(a(start));
      ++start;
      ++end;
      if (end >= 0 && end < n) sum.add(a(end));
    }
  }
}

// sliding_mean whose first window is assembled from block sums: each thread
// sums its own kPer values (stored to bs[] in LDS as float), and the first
// window of a thread adds the blocks it covers whole and reads only its two
// partial ends (about win/kPer + 2 kPer values instead of win).  The window
// sum then carries the float rounding of its whole blocks (relative 6e-8
// each); the slide itself runs in Acc.  Every thread must call it.
template <int kPer, class Acc, class A, class U>
__device__ __forceinline__ void sliding_mean_blocked(int n, int win, const A &a, const U &use,
                                                     float *bs) {
  const int t0 = (int)threadIdx.x * kPer;
  {
    Acc b;
#pragma unroll
    for (int m = 0; m < kPer; ++m)
      if (t0 + m < n) b.add(a(t0 + m));
    bs[threadIdx.x] = b.total();
  }
  __syncthreads();
  if (t0 >= n) return;
  int start = t0 - win / 2, end = start + win - 1;
  const int lo = max(start, 0), hi = min(end, n - 1);
  const int ub = (lo + kPer - 1) / kPer;  // first whole block
  const int ue = (hi + 1) / kPer;         // blocks [ub, ue) lie inside [lo, hi]
  Acc sum;
  if (ub < ue) {
    for (int i = lo; i < ub * kPer; ++i) sum.add(a(i));
    for (int u = ub; u < ue; ++u) sum.add(bs[u]);
    for (int i = ue * kPer; i <= hi; ++i) sum.add(a(i));
  } else {
    for (int i = lo; i <= hi; ++i) sum.add(a(i));
  }
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    const int t = t0 + m;
    if (t < n) {
      float w = sum.mean(win);
      if (!isfinite(w) || w < kMinVal) w = kMinVal;
      use(m, t, w);
      if (start >= 0 && start < n) sum.sub(a(start));
      ++start;
      ++end;
      if (end >= 0 && end < n) sum.add(a(end));
    }
  }
}

// sliding_mean of |X[k]| over kWin bins for the kFusedPerB outputs of each
// thread, with every |X| formed once: a thread forms its own kFusedPerB
// values and takes those of the two threads on each side from their lanes
// (__shfl), or, across a wave boundary, from xl (lanes 62, 63 of every wave)
// and xr (lanes 0, 1), 16 x 2 x kFusedPerB floats of LDS each.  Values
// outside [0, nbin) are 0, which the sums add exactly, in the same order as
// sliding_mean, so the weights are identical to it.  Every thread must call
// it (it holds a barrier).
template <int kWin, class Acc, class U>
__device__ __forceinline__ void whiten_mean_lanes(const float2 *X, int nbin, float *xl, float *xr,
                                                  const U &use) {
  constexpr int P = kFusedPerB;
  constexpr int H = kWin / 2;  // window of output t: bins [t - H, t - H + kWin - 1]
  static_assert(-H >= -2 * P && P - 1 - H + kWin < 3 * P, "neighbours within two threads");
  constexpr int kWaves = kS2sThreads / 64;
  const int T = threadIdx.x, lane = T & 63, wave = T >> 6;
  const int t0 = T * P;
  float own[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int k = t0 + m;
    float v = 0.f;
    if (k < nbin) {
      const float2 z = X[k];
      v = sqrtf(z.x * z.x + z.y * z.y);
    }
    own[m] = v;
  }
  if (lane >= 62) {
#pragma unroll
    for (int m = 0; m < P; ++m) xl[(wave * 2 + lane - 62) * P + m] = own[m];
  }
  if (lane < 2) {
#pragma unroll
    for (int m = 0; m < P; ++m) xr[(wave * 2 + lane) * P + m] = own[m];
  }
  __syncthreads();
  // |X| at bin t0 + r (r a constant after unrolling)
  auto val = [&](int r) -> float {
    const int d = r >= 0 ? r / P : -((P - 1 - r) / P);
    const int m = r - d * P;
    if (d == 0) return own[m];
    float v = __shfl(own[m], lane + d);
    const int l2 = lane + d;
    if (l2 < 0)
      v = wave > 0 ? xl[((wave - 1) * 2 + l2 + 2) * P + m] : 0.f;
    else if (l2 > 63)
      v = wave < kWaves - 1 ? xr[((wave + 1) * 2 + l2 - 64) * P + m] : 0.f;
    return v;
  };
  Acc sum;
#pragma unroll
  for (int r = -H; r <= kWin - 1 - H; ++r) sum.add(val(r));
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int t = t0 + m;
    const float vs = val(m - H), va = val(m - H + kWin);
    if (t < nbin) {
      float w = sum.mean(kWin);
      if (!isfinite(w) || w < kMinVal) w = kMinVal;
      use(m, t, w);
      sum.sub(vs);
      sum.add(va);
    }
  }
}

// Time-domain normalisation of the row x[0..n) in LDS (one channel):
// RUN-ABS x / w with w the running mean of |x| (0 for an invalid weight,
// non-finite -> 0, clip to +-MAXVAL), or one-bit sign(x)
// (normalization.cu:91-213, compute.cu:144-176).  bs: kS2sThreads floats.
template <class Acc>
__device__ __forceinline__ void normalise_row(float *x, int n, int mode, int win, float *bs) {
  if (mode == kLoadRunabs) {
    float y[kFusedPer];
    sliding_mean_blocked<kFusedPer, Acc>(n, win, [&](int i) { return fabsf(x[i]); },
                            [&](int m, int t, float d) {
                              float v = x[t];
                              v = (!isfinite(d) || d < kMinValD) ? 0.f : v / d;
                              if (!isfinite(v)) v = 0.f;
                              y[m] = v > kMaxVal ? kMaxVal : (v < -kMaxVal ? -kMaxVal : v);
                            },
                            bs);
    __syncthreads();
    const int t0 = (int)threadIdx.x * kFusedPer;
#pragma unroll
    for (int m = 0; m < kFusedPer; ++m)
      if (t0 + m < n) x[t0 + m] = y[m];
    __syncthreads();
  } else if (mode == kLoadOnebit) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const float v = x[i];
      x[i] = v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f);
    }
    __syncthreads();
  }
}

// Window cut + non-finite -> 0 + demean/detrend (FP64 sums) + cosine taper
// into the float view of LDS, zero padded to kTwoM samples
// (compute.cu:34-58, core/preprocess.cu:23-105).  Every load of the window
// is issued before the first LDS store, so their latencies overlap.
template <int kTwoM>
__device__ __forceinline__ void preprocess_row(float *ldsf, const float *__restrict__ src, long long off,
                                      int npts, int seg, int taper_size) {
  constexpr int kPer = (kTwoM + kS2sThreads - 1) / kS2sThreads;
  double s = 0.0, unused1 = 0.0, unused2 = 0.0;
  float v[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kS2sThreads;
    v[j] = (i < seg && off + i < npts) ? stream_load(src + off + i) : 0.f;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kS2sThreads;
    if (i < kTwoM) {
      const float x = isfinite(v[j]) ? v[j] : 0.f;
      ldsf[i] = x;
      s += x;  // zero past seg
    }
  }
  block_sum3(s, unused1, unused2);
  const double mean = s / seg;
  double s1 = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    const float v = (float)((double)ldsf[i] - mean);
    ldsf[i] = v;
    s1 += v;
    si += (double)v * i;
  }
  unused1 = 0.0;
  block_sum3(s1, si, unused1);
  const double W = seg;
  const double a12 = 0.5 * W * (W - 1);
  const double a11 = a12 * (2 * W - 1) / 3.;
  const double a22 = W;
  const double det = a11 * a22 - a12 * a12;
  const double ca = (a22 * si - a12 * s1) / det;
  const double cb = (a11 * s1 - a12 * si) / det;
  for (int i = threadIdx.x; i < seg; i += blockDim.x) {
    float v = (float)((double)ldsf[i] - (ca * i + cb));
    float factor = 1.f;
    if (i < taper_size)
      factor = (float)(0.5 * (1.0 - cos(M_PI * i / taper_size)));
    else if (i >= seg - taper_size)
      factor = (float)(0.5 * (1.0 - cos(M_PI * (seg - i) / taper_size)));
    ldsf[i] = v * factor;
  }
  __syncthreads();
}

// The last transform of a fused kernel: the real row x[0..seg) in LDS ->
// R2C(2 MN) x dt, sanitised, to X (whitening spectrum, k <= MN) or to the
// output spectrum (k < nspec, optional phase-only).
template <class FPN, bool kOut>
__device__ __forceinline__ void fused_emit(float2 *lds, int seg, const float2 *__restrict__ tw0n,
                                  const float2 *twln, const float2 *foldn, float dt,
                                  int phase_only, float2 *__restrict__ dst, int nspec) {
  float *ldsf = (float *)lds;
  if constexpr (kOut) {
    for (int i = seg + (int)threadIdx.x; i < 2 * FPN::kM; i += blockDim.x) ldsf[i] = 0.f;
    __syncthreads();
  }
  fixed_fft_fwd<FPN, kS2sThreads>(lds, tw0n, twln);
  fixed_r2c_emit<FPN>(lds, foldn, [&](int k, float2 v) {
    if (kOut && k >= nspec) return;
    v = cscale(v, dt);
    if (!cfinite(v)) v = make_float2(0.f, 0.f);
    if (kOut && phase_only) {
      const float a = hypotf(v.x, v.y);
      v = (!isfinite(a) || a <= kMinVal) ? make_float2(0.f, 0.f) : cscale(v, 1.f / a);
      if (!cfinite(v)) v = make_float2(0.f, 0.f);
    }
    stream_store(dst + k, v);
  }, (int)((reinterpret_cast<uintptr_t>(dst) >> 3) & 15));
}

template <class FPF, class FPN, bool kOut>
__global__ void __launch_bounds__(kS2sThreads)
s2s_front_kernel(const float *__restrict__ traces, int npts, int nfiles,
                 const int *__restrict__ steps, int shift, int seg, int taper_size,
                 FixedTables tf, FixedTables tn, const float *__restrict__ resp0, float dt,
                 double inv_weight, int norm_mode, int runabs_win, int phase_only, int comp_sums,
                 float2 *__restrict__ dst, long long dpitch, int nspec) {
  extern __shared__ float2 lds[];
  constexpr int kData = (FPF::kM > FPN::kM ? FPF::kM : FPN::kM) + 1;
  float2 *twlf = lds + kData;
  float2 *twln = twlf + fixed_lds_tw_count<FPF>();
  float2 *foldf = twln + fixed_lds_tw_count<FPN>();
  float2 *foldn = foldf + fold_lds_count(FPF::kM);
  float *bs = (float *)(foldn + fold_lds_count(FPN::kM));
  stage_fixed<FPF>(twlf, foldf, tf);
  stage_fixed<FPN>(twln, foldn, tn);
  float *ldsf = (float *)lds;
  const int row = blockIdx.x;
  const int step = row / nfiles, file = row - step * nfiles;
  preprocess_row<2 * FPF::kM>(ldsf, traces + (size_t)file * npts,
                              (long long)steps[step] * shift, npts, seg, taper_size);
  constexpr int Mf = FPF::kM;
  // this thread's response values, loaded while the forward transform runs
  float2 rsp[fixed_pair_trips<FPF>()];
#pragma unroll
  for (int it = 0; it < fixed_pair_trips<FPF>(); ++it) {
    const int k = (int)threadIdx.x + it * kS2sThreads;
    rsp[it] = make_float2(0.f, 0.f);
    if (2 * k <= Mf) rsp[it] = make_float2(resp0[k], resp0[k == 0 ? Mf : Mf - k]);
  }
  fixed_fft_fwd<FPF, kS2sThreads>(lds, tf.tw0, twlf);
  fixed_spectral_pairs<FPF>(lds, foldf,
                            [&](int it, int, float2 xk, float2 xm, float2 &yk, float2 &ym) {
                              xk = cscale(xk, dt);
                              xm = cscale(xm, dt);
                              if (!cfinite(xk)) xk = make_float2(0.f, 0.f);
                              if (!cfinite(xm)) xm = make_float2(0.f, 0.f);
                              yk = cscale(xk, rsp[it].x);
                              ym = cscale(xm, rsp[it].y);
                            });
  fixed_fft_inv<FPF, kS2sThreads>(lds, tf.tw0, twlf);
  for (int i = threadIdx.x; i < seg; i += blockDim.x) ldsf[i] = (float)((double)ldsf[i] * inv_weight);
  __syncthreads();
  if (comp_sums)
    normalise_row<AccComp>(ldsf, seg, norm_mode, runabs_win, bs);
  else
    normalise_row<Acc64>(ldsf, seg, norm_mode, runabs_win, bs);
  fused_emit<FPN, kOut>(lds, seg, tn.tw0, twln, foldn, dt, phase_only,
                        dst + (size_t)row * dpitch, nspec);
}

template <class FPS, class FPN, bool kOut>
__global__ void __launch_bounds__(kS2sThreads)
s2s_back_kernel(float2 *X, long long fpitch, int seg, int i1, int i2, int i3, int i4,
                int whiten_win, const float *__restrict__ staper, FixedTables ts, FixedTables tn,
                double inv_weight_seg,
                int norm_mode, int runabs_win, float dt, int phase_only, int comp_sums,
                float2 *dst, long long dpitch, int nspec) {
  extern __shared__ float2 lds[];
  constexpr int kData = (FPS::kM > FPN::kM ? FPS::kM : FPN::kM) + 1;
  constexpr int Ms = FPS::kM;
  float2 *twls = lds + kData;
  float2 *twln = twls + fixed_lds_tw_count<FPS>();
  float2 *folds = twln + fixed_lds_tw_count<FPN>();
  float2 *foldn = folds + fold_lds_count(FPS::kM);
  float *bs = (float *)(foldn + fold_lds_count(FPN::kM));
  stage_fixed<FPS>(twls, folds, ts);
  stage_fixed<FPN>(twln, foldn, tn);
  const int row = blockIdx.x;
  const float2 *Xr = X + (size_t)row * fpitch;
  {
    // every load of the row before the first LDS store
    constexpr int kPer = (Ms + 1 + kS2sThreads - 1) / kS2sThreads;
    float2 v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int k = threadIdx.x + j * kS2sThreads;
      if (k <= Ms) v[j] = stream_load(Xr + k);
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int k = threadIdx.x + j * kS2sThreads;
      if (k <= Ms) lds[k] = v[j];
    }
  }
  __syncthreads();
  // whitening weights (running mean of |X| over the bins, one channel) and
  // X / w x taper (staper[k] = spec_taper(k, i1, i2, i3, min(i4, seg/2+1)),
  // a table the plan fills), sanitised (spectrum.cu:205-300)
  const int nbin = Ms + 1;
  float wk[kFusedPerB];
  auto absx = [&](int k) {
    const float2 z = lds[k];
    return sqrtf(z.x * z.x + z.y * z.y);
  };
  auto keep = [&](int m, int, float d) { wk[m] = d; };
  // the window of one-hour windows at dt 0.1 s (int(0.02 seg dt) = 72 bins)
  // forms each |X| once; other windows slide on sliding_mean
  float *xl = bs + 16 * 2 * kFusedPerB;  // xr = bs [0, 16 x 2 x kFusedPerB), xl after it
  if (whiten_win == 72 && comp_sums)
    whiten_mean_lanes<72, AccComp>(lds, nbin, xl, bs, keep);
  else if (whiten_win == 72)
    whiten_mean_lanes<72, Acc64>(lds, nbin, xl, bs, keep);
  else if (comp_sums)
    sliding_mean<kFusedPerB, AccComp>(nbin, whiten_win, absx, keep);
  else
    sliding_mean<kFusedPerB, Acc64>(nbin, whiten_win, absx, keep);
  __syncthreads();  // every |X| read is done: each thread rewrites its own bins
  {
    const int k0 = (int)threadIdx.x * kFusedPerB;
#pragma unroll
    for (int m = 0; m < kFusedPerB; ++m) {
      const int k = k0 + m;
      if (k < nbin) {
        const float d = wk[m];
        float2 v = lds[k];
        v = (!isfinite(d) || d < kMinValD) ? make_float2(0.f, 0.f) : make_float2(v.x / d, v.y / d);
        v = cscale(v, staper[k]);
        if (!cfinite(v)) v = make_float2(0.f, 0.f);
        lds[k] = v;
      }
    }
  }
  __syncthreads();
  // fold into the digit-reversed positions of the inverse (every value read
  // before any is written): Z[k] = S + i D with S = X[k] + conj X[Ms-k],
  // D = conj(w_k) (X[k] - conj X[Ms-k]), w_k = exp(-i pi k / Ms).  Bin Ms-k
  // folds from the same two values, Z[Ms-k] = conj S + i conj D (w_{Ms-k} =
  // -conj w_k), so a thread takes the pair (k, Ms-k) for k = 0 .. Ms/2; k = 0
  // (with X[Ms]) and k = Ms/2 fold alone.  Two loads per pair and 2 x
  // kPairPer values held across the barrier.
  static_assert(Ms % 2 == 0, "pairwise fold needs an even half length");
  constexpr int kHalf = Ms / 2;
  constexpr int kPairPer = (kHalf + 1 + kS2sThreads - 1) / kS2sThreads;
  float2 zk[kPairPer], zm[kPairPer];
#pragma unroll
  for (int i = 0; i < kPairPer; ++i) {
    const int k = threadIdx.x + i * kS2sThreads;
    if (k <= kHalf) {
      float2 a = lds[k];
      float2 b = cconj(lds[Ms - k]);
      if (k == 0) {
        a.y = 0.f;
        b.y = 0.f;
      }
      const float2 sv = cadd(a, b);
      const float2 dv = cmul(cconj(fold_lookup(folds, k)), csub(a, b));
      zk[i] = make_float2(sv.x - dv.y, sv.y + dv.x);
      zm[i] = make_float2(sv.x + dv.y, dv.x - sv.y);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPairPer; ++i) {
    const int k = threadIdx.x + i * kS2sThreads;
    if (k <= kHalf) {
      lds[fixed_pos<FPS>(k)] = zk[i];
      if (k != 0 && k != kHalf) lds[fixed_pos<FPS>(Ms - k)] = zm[i];
    }
  }
  __syncthreads();
  fixed_fft_inv<FPS, kS2sThreads>(lds, ts.tw0, twls);
  float *ldsf = (float *)lds;
  for (int i = threadIdx.x; i < seg; i += blockDim.x)
    ldsf[i] = (float)((double)ldsf[i] * inv_weight_seg);
  __syncthreads();
  if (comp_sums)
    normalise_row<AccComp>(ldsf, seg, norm_mode, runabs_win, bs);
  else
    normalise_row<Acc64>(ldsf, seg, norm_mode, runabs_win, bs);
  fused_emit<FPN, kOut>(lds, seg, tn.tw0, twln, foldn, dt, phase_only, dst + (size_t)row * dpitch,
                        nspec);
}

}  // namespace fxc

using namespace fxc;

// ---------------------------------------------------------------------------
// Plan object behind the C ABI.
// ---------------------------------------------------------------------------
struct fxc_s2s_plan {
  fxc_s2s_config cfg{};
  int nspec = 0;
  int fwidth = 0;          // seg/2 + 1
  long long fpitch = 0;    // complex row pitch of whitening spectra
  long long bpitch = 0;    // complex row pitch of the MF base spectrum
  int taper_size = 0;
  int runabs_win = 0;
  int whiten_win = 0;
  std::vector<int> mf_win;
  float *d_resp = nullptr;  // [filter_count][nfft_filter/2 + 1]
  RealTables tf{}, ts{}, to{};
  // workspace (grows on demand)
  long long cap_rows = 0, cap_frames = 0;
  float *d_x = nullptr, *d_w = nullptr, *d_xb = nullptr, *d_total = nullptr;
  float2 *d_X = nullptr, *d_base = nullptr;
  float *d_amp = nullptr, *d_ww = nullptr;
  int *d_steps = nullptr;
  int cap_steps = 0;
  // long-window path (transforms beyond one workgroup's LDS)
  bool large = false;
  long long lpitch = 0;              // complex row pitch of the long spectra
  float2 *d_L = nullptr;             // [rows][lpitch]
  float2 *d_Lw = nullptr;            // four-step work, kLargeChunk rows
  long long cap_L = 0;
  // fused one-channel path at the production lengths (s2s_front/back_kernel)
  bool fused = false;
  float2 *d_fx = nullptr;  // pass-0 and LDS base twiddles of the three plans
  float *d_staper = nullptr;  // spectral taper factor of each whitening bin
  int comp_sums = 0;          // spec mode: compensated-FP32 running means (fxc_s2s_set_sums)
  FixedTables fxf{}, fxs{}, fxo{};
};

// Rows per four-step call of the long-window path (bounds its work buffer).
constexpr long long kLargeChunk = 2048;

static int make_tables(int n, RealTables *t) {
  const RealFftPlan *f = nullptr, *i = nullptr;
  FXC_TRY(plan_cache_real(n, -1, &f));
  FXC_TRY(plan_cache_real(n, 1, &i));
  t->plan = f->half.dev;
  t->tw_fwd = f->half.d_tw;
  t->tw_inv = i->half.d_tw;
  t->fold = f->d_fold;
  return FXC_OK;
}

static size_t lds_bytes(int data_slots, int M) {
  return ((size_t)data_slots + tables_count(M)) * sizeof(float2);
}

template <class K>
static int set_lds(K kern, size_t bytes) {
  FXC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes));
  return FXC_OK;
}

// The fused path applies to one channel, every normalisation but RUN-ABS-MF,
// and the production lengths; FXC_S2S_FUSED=0 keeps the five-kernel path.
static int s2s_fused_setup(fxc_s2s_plan *p) {
  const fxc_s2s_config &c = p->cfg;
  const char *env = getenv("FXC_S2S_FUSED");
  if (env && atoi(env) == 0) return FXC_OK;
  if (c.num_ch != 1 || c.normalize == 1 || c.nfft_filter / 2 != S2sFixF::kM ||
      c.seg / 2 != S2sFixS::kM || c.nfft_out / 2 != S2sFixO::kM)
    return FXC_OK;
  std::vector<float2> all;
  size_t off[6];
  auto add = [&](const std::vector<float2> &v, size_t *o) {
    *o = all.size();
    all.insert(all.end(), v.begin(), v.end());
  };
  add(fixed_pass0_twiddles<S2sFixF>(), &off[0]);
  add(fixed_lds_twiddles<S2sFixF>(-1), &off[1]);
  add(fixed_pass0_twiddles<S2sFixS>(), &off[2]);
  add(fixed_lds_twiddles<S2sFixS>(-1), &off[3]);
  add(fixed_pass0_twiddles<S2sFixO>(), &off[4]);
  add(fixed_lds_twiddles<S2sFixO>(-1), &off[5]);
  FXC_TRY(dmalloc((void **)&p->d_fx, all.size() * sizeof(float2)));
  FXC_HIP(hipMemcpy(p->d_fx, all.data(), all.size() * sizeof(float2), hipMemcpyHostToDevice));
  p->fxf = {p->d_fx + off[0], p->d_fx + off[1], p->tf.fold};
  p->fxs = {p->d_fx + off[2], p->d_fx + off[3], p->ts.fold};
  p->fxo = {p->d_fx + off[4], p->d_fx + off[5], p->to.fold};
  // the spectral taper of the whitening bins, as spec_taper forms it
  // (double cosine, rounded to float) -- a table instead of a cosine per bin
  {
    const int nbin = c.seg / 2 + 1;
    const int i1 = c.idx[0], i2 = c.idx[1], i3 = c.idx[2], i4 = std::min(c.idx[3], nbin);
    std::vector<float> tp((size_t)nbin);
    for (int k = 0; k < nbin; ++k) {
      float f = 1.f;
      if (k < i1) f = 0.f;
      else if (k < i2) f = (float)((1.0 - std::cos(M_PI / (i2 - i1) * (k - i1))) / 2.0);
      else if (k >= i3 && k < i4) f = (float)((1.0 + std::cos(M_PI / (i4 - i3) * (k - i3))) / 2.0);
      else if (k >= i4) f = 0.f;
      tp[(size_t)k] = f;
    }
    FXC_TRY(dmalloc((void **)&p->d_staper, tp.size() * sizeof(float)));
    FXC_HIP(hipMemcpy(p->d_staper, tp.data(), tp.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  p->fused = true;
  return FXC_OK;
}

extern "C" int fxc_s2s_plan_create(fxc_s2s_plan **out, const fxc_s2s_config *cfg) {
  FXC_CHECK_ARG(out && cfg, "fxc_s2s_plan_create: null argument");
  *out = nullptr;
  const fxc_s2s_config &c = *cfg;
  FXC_CHECK_ARG(c.npts >= 1 && c.seg >= 2 && c.seg % 2 == 0 && c.shift >= 1 && c.num_ch >= 1,
                "fxc_s2s_plan_create: invalid lengths npts=%d seg=%d shift=%d nch=%d", c.npts,
                c.seg, c.shift, c.num_ch);
  FXC_CHECK_ARG(c.nfft_filter >= c.seg && c.nfft_filter % 2 == 0 && c.nfft_out >= c.seg &&
                    c.nfft_out % 2 == 0,
                "fxc_s2s_plan_create: FFT lengths must be even and >= seg");
  FXC_CHECK_ARG(c.delta > 0.f && c.freq_low > 0.f, "fxc_s2s_plan_create: invalid delta/freq_low");
  FXC_CHECK_ARG(c.whiten >= 0 && c.whiten <= 3 && c.normalize >= 0 && c.normalize <= 3,
                "fxc_s2s_plan_create: whiten must be 0..3 and normalize 0..3");
  FXC_CHECK_ARG(c.filter_count >= 1 && c.responses && c.freq_lows,
                "fxc_s2s_plan_create: at least one filter response is required");
  FXC_CHECK_ARG(c.normalize != 1 || c.filter_count >= 2,
                "fxc_s2s_plan_create: RUN-ABS-MF needs sub-band filters");
  bool large = false;
  {
    // exact LDS needs of the per-row kernels (data slots + staged tables)
    const int Mf = c.nfft_filter / 2, Ms = c.seg / 2, Mo = c.nfft_out / 2;
    const int need_f = Mf + 1 + tables_count(Mf);
    const int need_s = Ms + 1 + tables_count(Ms);
    const int need_o = Mo + 1 + tables_count(Mo);
    const int need_so = std::max(Ms, Mo) + 1 + tables_count(Ms) + tables_count(Mo);
    const int need = std::max(std::max(need_f, need_s), std::max(need_o, need_so));
    large = need > kLdsComplexCapacity;
    if (large) {
      // every transform of the plan runs as a four-step transform over HBM
      int a = 0, b = 0;
      for (int M : {Mf, Ms, Mo})
        FXC_CHECK_ARG(large_factor(M, &a, &b),
                      "fxc_s2s_plan_create: transforms (seg %d, filter %d, out %d) exceed one "
                      "workgroup and %d has no two-level factorisation", c.seg, c.nfft_filter,
                      c.nfft_out, 2 * M);
    }
  }
  fxc_s2s_plan *p = new fxc_s2s_plan();
  p->large = large;
  p->lpitch = ((long long)std::max(std::max(c.nfft_filter, c.seg), c.nfft_out) / 2 + 1 + 15) /
              16 * 16;
  p->cfg = c;
  p->cfg.responses = nullptr;
  p->cfg.freq_lows = nullptr;
  p->nspec = c.nfft_out / 2 + 1;
  p->fwidth = c.seg / 2 + 1;
  p->fpitch = (p->fwidth + 15) / 16 * 16;
  p->bpitch = ((c.nfft_filter / 2 + 1) + 15) / 16 * 16;
  // taper length (size_t)(2.0f*(1.0f/freq_low)/delta), preprocess.cu:91
  p->taper_size = (int)(2.0f * (1.0f / c.freq_low) / c.delta);
  // RUN-ABS window 2*int(1.0/(f*dt))+1 with f = freq_low*0.667f (compute.cu:173,
  // normalization.cu:126)
  const float fra = c.freq_low * 0.667f;
  p->runabs_win = std::max(2 * (int)(1.0 / (double)(fra * c.delta)) + 1, 1);
  for (int i = 0; i < c.filter_count; ++i)
    p->mf_win.push_back(std::max(2 * (int)(1.0 / (double)(c.freq_lows[i] * c.delta)) + 1, 1));
  // whitening smoother int(0.02*pitch*delta) (spectrum.cu:205)
  p->whiten_win = std::max((int)(0.02 * c.seg * (double)c.delta), 1);
  int rc = FXC_OK;
  const size_t rbytes = (size_t)c.filter_count * (c.nfft_filter / 2 + 1) * sizeof(float);
  if ((rc = dmalloc((void **)&p->d_resp, rbytes)) != FXC_OK) {
    delete p;
    return rc;
  }
  if (hipMemcpy(p->d_resp, c.responses, rbytes, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("fxc_s2s_plan_create: response upload failed");
    dfree(p->d_resp);
    delete p;
    return FXC_ERR_HIP;
  }
  if (!large && ((rc = make_tables(c.nfft_filter, &p->tf)) != FXC_OK ||
                 (rc = make_tables(c.seg, &p->ts)) != FXC_OK ||
                 (rc = make_tables(c.nfft_out, &p->to)) != FXC_OK)) {
    dfree(p->d_resp);
    delete p;
    return rc;
  }
  if (!large && (rc = s2s_fused_setup(p)) != FXC_OK) {
    fxc_s2s_plan_destroy(p);
    return rc;
  }
  *out = p;
  return FXC_OK;
}

extern "C" int fxc_s2s_plan_destroy(fxc_s2s_plan *p) {
  if (!p) return FXC_OK;
  for (void *q : {(void *)p->d_resp, (void *)p->d_x, (void *)p->d_w, (void *)p->d_xb,
                  (void *)p->d_total, (void *)p->d_X, (void *)p->d_base, (void *)p->d_amp,
                  (void *)p->d_ww, (void *)p->d_steps, (void *)p->d_L, (void *)p->d_Lw,
                  (void *)p->d_fx, (void *)p->d_staper})
    dfree(q);
  delete p;
  return FXC_OK;
}

extern "C" int fxc_s2s_set_sums(fxc_s2s_plan *p, int mode) {
  FXC_CHECK_ARG(p && (mode == 0 || mode == 1), "fxc_s2s_set_sums: invalid arguments");
  if (mode == 1 && !p->fused) {
    set_error("fxc_s2s_set_sums: compensated-FP32 running means need the fused one-channel path");
    return FXC_ERR_UNSUPPORTED;
  }
  p->comp_sums = mode;
  return FXC_OK;
}

extern "C" int fxc_s2s_plan_info(const fxc_s2s_plan *p, int *nspec, int *taper_size,
                                 int *runabs_win, int *whiten_win) {
  FXC_CHECK_ARG(p, "fxc_s2s_plan_info: null plan");
  if (nspec) *nspec = p->nspec;
  if (taper_size) *taper_size = p->taper_size;
  if (runabs_win) *runabs_win = p->runabs_win;
  if (whiten_win) *whiten_win = p->whiten_win;
  return FXC_OK;
}

static int s2s_ensure(fxc_s2s_plan *p, long long rows, long long frames, int nsteps) {
  const fxc_s2s_config &c = p->cfg;
  if (p->fused) {
    // only the whitening spectra X travel between the fused kernels
    if (c.whiten && rows > p->cap_rows) {
      dfree(p->d_X);
      p->d_X = nullptr;
      p->cap_rows = 0;
      FXC_TRY(dmalloc((void **)&p->d_X, (size_t)rows * p->fpitch * sizeof(float2)));
      p->cap_rows = rows;
    }
    frames = 0;
    rows = 0;
  }
  if (rows > p->cap_rows) {
    for (void *q : {(void *)p->d_x, (void *)p->d_xb, (void *)p->d_total, (void *)p->d_X,
                    (void *)p->d_base, (void *)p->d_amp})
      dfree(q);
    p->d_x = p->d_xb = p->d_total = p->d_amp = nullptr;
    p->d_X = p->d_base = nullptr;
    p->cap_rows = 0;
    FXC_TRY(dmalloc((void **)&p->d_x, (size_t)rows * c.seg * sizeof(float)));
    if (c.whiten) {
      FXC_TRY(dmalloc((void **)&p->d_X, (size_t)rows * p->fpitch * sizeof(float2)));
      FXC_TRY(dmalloc((void **)&p->d_amp, (size_t)rows * p->fpitch * sizeof(float)));
    }
    if (c.normalize == 1) {
      FXC_TRY(dmalloc((void **)&p->d_xb, (size_t)rows * c.seg * sizeof(float)));
      FXC_TRY(dmalloc((void **)&p->d_total, (size_t)rows * c.seg * sizeof(float)));
      FXC_TRY(dmalloc((void **)&p->d_base, (size_t)rows * p->bpitch * sizeof(float2)));
    }
    p->cap_rows = rows;
  }
  if (frames > p->cap_frames) {
    dfree(p->d_w);
    dfree(p->d_ww);
    p->d_w = p->d_ww = nullptr;
    p->cap_frames = 0;
    FXC_TRY(dmalloc((void **)&p->d_w, (size_t)frames * c.seg * sizeof(float)));
    if (c.whiten) FXC_TRY(dmalloc((void **)&p->d_ww, (size_t)frames * p->fpitch * sizeof(float)));
    p->cap_frames = frames;
  }
  if (nsteps > p->cap_steps) {
    dfree(p->d_steps);
    p->d_steps = nullptr;
    FXC_TRY(dmalloc((void **)&p->d_steps, (size_t)nsteps * sizeof(int)));
    p->cap_steps = nsteps;
  }
  return FXC_OK;
}

static int launch_weight(bool cplx, const void *data, long long pitch, int width, int nch, int win,
                         long long frames, float *w, long long wpitch, hipStream_t st) {
  if (win >= 1 && win <= kWTileMaxWin) {
    const unsigned tiles = (unsigned)ceil_div((size_t)width, (size_t)kWTile);
    const size_t lds = (size_t)(kWTile + win - 1) * sizeof(double);
    const dim3 grid((unsigned)frames, tiles);
    const void *fn = cplx ? (const void *)s2s_weight_tile_kernel<true>
                          : (const void *)s2s_weight_tile_kernel<false>;
    FXC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (cplx)
      hipLaunchKernelGGL(s2s_weight_tile_kernel<true>, grid, dim3(kWThreads), lds, st, data, pitch,
                         width, nch, win, w, wpitch);
    else
      hipLaunchKernelGGL(s2s_weight_tile_kernel<false>, grid, dim3(kWThreads), lds, st, data,
                         pitch, width, nch, win, w, wpitch);
    FXC_HIP(hipGetLastError());
    return FXC_OK;
  }
  // very long windows (bands far below 0.001 cycles per sample): rolling form
  FXC_CHECK_ARG(frames <= 65535, "sac2spec: %lld frames exceed one launch", frames);
  const int per_block = 256 * kWeightChunk;
  const unsigned gx = (unsigned)ceil_div((size_t)width, (size_t)per_block);
  const dim3 grid(gx, (unsigned)frames);
  if (cplx)
    hipLaunchKernelGGL(s2s_weight_kernel<true>, grid, dim3(256), 0, st, data, pitch, width, nch,
                       win, w, wpitch);
  else
    hipLaunchKernelGGL(s2s_weight_kernel<false>, grid, dim3(256), 0, st, data, pitch, width, nch,
                       win, w, wpitch);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

// Whitening of the rows in d_x (mode/w: pending normalisation).  With
// `out` set the inverse also produces the output spectrum.
static int s2s_whiten(fxc_s2s_plan *p, long long rows, long long frames, int mode, const float *w,
                      float2 *out, hipStream_t st) {
  const fxc_s2s_config &c = p->cfg;
  const int Ms = c.seg / 2;
  size_t lds = lds_bytes(Ms + 1, Ms);
  FXC_TRY(set_lds(s2s_whiten_fwd_kernel, lds));
  hipLaunchKernelGGL(s2s_whiten_fwd_kernel, dim3((unsigned)rows), dim3(kS2sThreads), lds, st,
                     p->d_x, c.seg, c.num_ch, mode, w, p->ts, c.delta, p->d_X, p->d_amp, p->fpitch);
  FXC_HIP(hipGetLastError());
  FXC_TRY(launch_weight(true, p->d_X, p->fpitch, p->fwidth, c.num_ch, p->whiten_win, frames,
                        p->d_ww, p->fpitch, st));
  const int Mo = c.nfft_out / 2;
  const int data = std::max(Ms, out ? Mo : Ms) + 1;
  lds = ((size_t)data + tables_count(Ms) + (out ? tables_count(Mo) : 0)) * sizeof(float2);
  FXC_CHECK_ARG(lds <= (size_t)kLdsComplexCapacity * sizeof(float2),
                "whitening + output transform exceed LDS");
  FXC_TRY(set_lds(s2s_whiten_inv_kernel, lds));
  const double inv_seg = 1.0 / (double)((float)c.seg * c.delta);
  hipLaunchKernelGGL(s2s_whiten_inv_kernel, dim3((unsigned)rows), dim3(kS2sThreads), lds, st,
                     p->d_X, p->fpitch, p->d_ww, c.num_ch, c.seg, c.idx[0], c.idx[1], c.idx[2],
                     c.idx[3], p->ts, inv_seg, p->d_x, p->to, c.delta, c.phase_only, out,
                     p->nspec);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

// Long-window path: four-step real transforms in chunks of kLargeChunk rows.
static int s2sL_r2c(fxc_s2s_plan *p, int n, const float *x, long long xd, int in_len, float2 *X,
                    long long Xd, long long rows, hipStream_t st) {
  for (long long r = 0; r < rows; r += kLargeChunk)
    FXC_TRY(large_r2c(n, x + r * xd, xd, in_len, X + r * Xd, Xd, std::min(kLargeChunk, rows - r),
                      p->d_Lw, 1.f, st));
  return FXC_OK;
}

static int s2sL_c2r(fxc_s2s_plan *p, int n, const float2 *X, long long Xd, float *x, long long xd,
                    int out_len, double scale, long long rows, hipStream_t st) {
  for (long long r = 0; r < rows; r += kLargeChunk)
    FXC_TRY(large_c2r(n, X + r * Xd, Xd, x + r * xd, xd, out_len, std::min(kLargeChunk, rows - r),
                      p->d_Lw, scale, st));
  return FXC_OK;
}

static dim3 s2sL_grid(long long n, long long rows) {
  return dim3((unsigned)ceil_div((size_t)n, (size_t)256), (unsigned)rows);
}

// Whitening of d_x rows (pending normalisation `mode`, weights w first).
static int s2sL_whiten(fxc_s2s_plan *p, long long rows, long long frames, int mode,
                       const float *w, hipStream_t st) {
  const fxc_s2s_config &c = p->cfg;
  if (mode != kLoadPlain) {
    hipLaunchKernelGGL(s2sL_norm_kernel, s2sL_grid(c.seg, rows), dim3(256), 0, st, p->d_x, c.seg,
                       c.num_ch, mode, w);
    FXC_HIP(hipGetLastError());
  }
  FXC_TRY(s2sL_r2c(p, c.seg, p->d_x, c.seg, c.seg, p->d_X, p->fpitch, rows, st));
  hipLaunchKernelGGL(s2sL_spec_kernel, s2sL_grid(p->fwidth, rows), dim3(256), 0, st, p->d_X,
                     p->fpitch, p->fwidth, c.delta, (const float *)nullptr, p->d_amp, p->fpitch);
  FXC_HIP(hipGetLastError());
  FXC_TRY(launch_weight(true, p->d_X, p->fpitch, p->fwidth, c.num_ch, p->whiten_win, frames,
                        p->d_ww, p->fpitch, st));
  hipLaunchKernelGGL(s2sL_whiten_kernel, s2sL_grid(p->fwidth, rows), dim3(256), 0, st, p->d_X,
                     p->fpitch, p->fwidth, p->d_ww, c.num_ch, c.idx[0], c.idx[1], c.idx[2],
                     c.idx[3]);
  FXC_HIP(hipGetLastError());
  const double inv_seg = 1.0 / (double)((float)c.seg * c.delta);
  return s2sL_c2r(p, c.seg, p->d_X, p->fpitch, p->d_x, c.seg, c.seg, inv_seg, rows, st);
}

static int s2s_process_large(fxc_s2s_plan *p, const float *d_traces, int nfiles, long long rows,
                             long long frames, float2 *d_out, hipStream_t st) {
  const fxc_s2s_config &c = p->cfg;
  FXC_CHECK_ARG(rows <= 65535, "sac2spec: %lld station-windows exceed one long-window call "
                "(65535)", rows);
  if (rows > p->cap_L) {
    dfree(p->d_L);
    dfree(p->d_Lw);
    p->d_L = p->d_Lw = nullptr;
    p->cap_L = 0;
    const int Mmax = std::max(std::max(c.nfft_filter, c.seg), c.nfft_out) / 2;
    FXC_TRY(dmalloc((void **)&p->d_L, (size_t)rows * p->lpitch * sizeof(float2)));
    FXC_TRY(dmalloc((void **)&p->d_Lw, (size_t)std::min(kLargeChunk, rows) * 2 * Mmax *
                                           sizeof(float2)));
    p->cap_L = rows;
  }
  // window cut + preprocess, then the broad-band prefilter
  hipLaunchKernelGGL(s2sL_pre_kernel, dim3((unsigned)rows), dim3(kS2sThreads), 0, st, d_traces,
                     c.npts, nfiles, p->d_steps, c.shift, c.seg, p->taper_size, p->d_x);
  FXC_HIP(hipGetLastError());
  const int nbf = c.nfft_filter / 2 + 1;
  const double inv_f = 1.0 / (double)((float)c.nfft_filter * c.delta);
  FXC_TRY(s2sL_r2c(p, c.nfft_filter, p->d_x, c.seg, c.seg, p->d_L, p->lpitch, rows, st));
  hipLaunchKernelGGL(s2sL_spec_kernel, s2sL_grid(nbf, rows), dim3(256), 0, st, p->d_L, p->lpitch,
                     nbf, c.delta, p->d_resp, (float *)nullptr, 0LL);
  FXC_HIP(hipGetLastError());
  FXC_TRY(s2sL_c2r(p, c.nfft_filter, p->d_L, p->lpitch, p->d_x, c.seg, c.seg, inv_f, rows, st));
  int mode = kLoadPlain;
  const float *w = nullptr;
  if (c.whiten == 1 || c.whiten == 3) FXC_TRY(s2sL_whiten(p, rows, frames, kLoadPlain, nullptr, st));
  if (c.normalize == 3) {
    FXC_TRY(launch_weight(false, p->d_x, c.seg, c.seg, c.num_ch, p->runabs_win, frames, p->d_w,
                          c.seg, st));
    mode = kLoadRunabs;
    w = p->d_w;
  } else if (c.normalize == 2) {
    mode = kLoadOnebit;
  } else if (c.normalize == 1) {
    // base spectrum R2C(nfft_filter) x dt, then per band C2R(base |H_i|^2)
    FXC_TRY(s2sL_r2c(p, c.nfft_filter, p->d_x, c.seg, c.seg, p->d_base, p->bpitch, rows, st));
    hipLaunchKernelGGL(s2sL_spec_kernel, s2sL_grid(nbf, rows), dim3(256), 0, st, p->d_base,
                       p->bpitch, nbf, c.delta, (const float *)nullptr, (float *)nullptr, 0LL);
    FXC_HIP(hipGetLastError());
    FXC_HIP(hipMemsetAsync(p->d_total, 0, (size_t)rows * c.seg * sizeof(float), st));
    for (int i = 1; i < c.filter_count; ++i) {
      hipLaunchKernelGGL(s2sL_band_kernel, s2sL_grid(nbf, rows), dim3(256), 0, st, p->d_base,
                         p->bpitch, nbf, p->d_resp + (size_t)i * nbf, p->d_L);
      FXC_HIP(hipGetLastError());
      FXC_TRY(s2sL_c2r(p, c.nfft_filter, p->d_L, p->bpitch, p->d_xb, c.seg, c.seg, inv_f, rows,
                       st));
      FXC_TRY(launch_weight(false, p->d_xb, c.seg, c.seg, c.num_ch, p->mf_win[i], frames, p->d_w,
                            c.seg, st));
      const int nbr = (int)ceil_div(c.seg, 256);
      hipLaunchKernelGGL(s2s_mf_accum_kernel, dim3((unsigned)(rows * nbr)), dim3(256), 0, st,
                         p->d_xb, p->d_w, c.seg, c.num_ch, nbr, p->d_total);
      FXC_HIP(hipGetLastError());
    }
    FXC_HIP(hipMemcpyAsync(p->d_x, p->d_total, (size_t)rows * c.seg * sizeof(float),
                           hipMemcpyDeviceToDevice, st));
  }
  if (c.whiten == 2 || c.whiten == 3) {
    FXC_TRY(s2sL_whiten(p, rows, frames, mode, w, st));
  } else if (mode != kLoadPlain) {
    hipLaunchKernelGGL(s2sL_norm_kernel, s2sL_grid(c.seg, rows), dim3(256), 0, st, p->d_x, c.seg,
                       c.num_ch, mode, w);
    FXC_HIP(hipGetLastError());
  }
  // output spectrum: zero pad to nfft_out, R2C, x dt, sanitise, phase-only
  FXC_TRY(s2sL_r2c(p, c.nfft_out, p->d_x, c.seg, c.seg, d_out, p->nspec, rows, st));
  hipLaunchKernelGGL(s2sL_out_kernel, s2sL_grid(p->nspec, rows), dim3(256), 0, st, d_out,
                     p->nspec, c.delta, c.phase_only);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

template <class FPN, bool kOut>
static int launch_front(fxc_s2s_plan *p, const float *d_traces, int nfiles, long long rows,
                        const FixedTables &tn, int norm, float2 *dst, long long dpitch,
                        hipStream_t st) {
  const fxc_s2s_config &c = p->cfg;
  constexpr size_t lds = (size_t)fused_lds_count<S2sFixF, FPN>() * sizeof(float2);
  static_assert(lds <= (size_t)kLdsComplexCapacity * sizeof(float2), "fused front fits LDS");
  auto kern = s2s_front_kernel<S2sFixF, FPN, kOut>;
  FXC_TRY(set_lds(kern, lds));
  const double inv_f = 1.0 / (double)((float)c.nfft_filter * c.delta);
  hipLaunchKernelGGL(kern, dim3((unsigned)rows), dim3(kS2sThreads), lds, st, d_traces, c.npts,
                     nfiles, p->d_steps, c.shift, c.seg, p->taper_size, p->fxf, tn, p->d_resp,
                     c.delta, inv_f, norm, p->runabs_win, c.phase_only, p->comp_sums, dst, dpitch,
                     p->nspec);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

template <class FPN, bool kOut>
static int launch_back(fxc_s2s_plan *p, long long rows, const FixedTables &tn, int norm,
                       float2 *dst, long long dpitch, hipStream_t st) {
  const fxc_s2s_config &c = p->cfg;
  constexpr size_t lds = (size_t)fused_lds_count<S2sFixS, FPN>() * sizeof(float2);
  static_assert(lds <= (size_t)kLdsComplexCapacity * sizeof(float2), "fused back fits LDS");
  auto kern = s2s_back_kernel<S2sFixS, FPN, kOut>;
  FXC_TRY(set_lds(kern, lds));
  const double inv_s = 1.0 / (double)((float)c.seg * c.delta);
  hipLaunchKernelGGL(kern, dim3((unsigned)rows), dim3(kS2sThreads), lds, st, p->d_X, p->fpitch,
                     c.seg, c.idx[0], c.idx[1], c.idx[2], c.idx[3], p->whiten_win, p->d_staper,
                     p->fxs, tn, inv_s, norm, p->runabs_win, c.delta, c.phase_only, p->comp_sums,
                     dst, dpitch, p->nspec);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

// Fused one-channel path: whiten 0 -> front(norm) to the output; 2 ->
// front(norm) to X, back to the output; 1 -> front to X, back(norm) to the
// output; 3 -> front to X, back(norm) to X, back to the output.
static int s2s_process_fused(fxc_s2s_plan *p, const float *d_traces, int nfiles, long long rows,
                             float2 *d_out, hipStream_t st) {
  const fxc_s2s_config &c = p->cfg;
  FXC_CHECK_ARG(rows <= 0x7fffffffLL, "sac2spec: %lld rows exceed one launch", rows);
  const int norm = c.normalize == 3 ? kLoadRunabs : (c.normalize == 2 ? kLoadOnebit : kLoadPlain);
  const long long np = p->nspec;
  switch (c.whiten) {
    case 0:
      return launch_front<S2sFixO, true>(p, d_traces, nfiles, rows, p->fxo, norm, d_out, np, st);
    case 2:
      FXC_TRY((launch_front<S2sFixS, false>(p, d_traces, nfiles, rows, p->fxs, norm, p->d_X,
                                           p->fpitch, st)));
      return launch_back<S2sFixO, true>(p, rows, p->fxo, kLoadPlain, d_out, np, st);
    case 1:
      FXC_TRY((launch_front<S2sFixS, false>(p, d_traces, nfiles, rows, p->fxs, kLoadPlain, p->d_X,
                                           p->fpitch, st)));
      return launch_back<S2sFixO, true>(p, rows, p->fxo, norm, d_out, np, st);
    default:
      FXC_TRY((launch_front<S2sFixS, false>(p, d_traces, nfiles, rows, p->fxs, kLoadPlain, p->d_X,
                                           p->fpitch, st)));
      FXC_TRY((launch_back<S2sFixS, false>(p, rows, p->fxs, norm, p->d_X, p->fpitch, st)));
      return launch_back<S2sFixO, true>(p, rows, p->fxo, kLoadPlain, d_out, np, st);
  }
}

// Processes nfiles traces [nfiles][npts] (nfiles = groups * num_ch) for the
// windows listed in h_steps; writes [nsteps][nfiles][nspec] complex64.
extern "C" int fxc_s2s_process(fxc_s2s_plan *p, const float *d_traces, int nfiles,
                               const int *h_steps, int nsteps, float *d_out, void *stream) {
  FXC_CHECK_ARG(p && d_traces && d_out && h_steps && nfiles >= 1 && nsteps >= 1,
                "fxc_s2s_process: invalid arguments");
  const fxc_s2s_config &c = p->cfg;
  FXC_CHECK_ARG(nfiles % c.num_ch == 0, "fxc_s2s_process: nfiles must be a multiple of num_ch");
  for (int i = 0; i < nsteps; ++i)
    FXC_CHECK_ARG(h_steps[i] >= 0, "fxc_s2s_process: negative step index");
  hipStream_t st = (hipStream_t)stream;
  const long long rows = (long long)nfiles * nsteps;
  const long long frames = rows / c.num_ch;
  FXC_TRY(s2s_ensure(p, rows, frames, nsteps));
  FXC_HIP(hipMemcpyAsync(p->d_steps, h_steps, nsteps * sizeof(int), hipMemcpyHostToDevice, st));
  if (p->large) return s2s_process_large(p, d_traces, nfiles, rows, frames, (float2 *)d_out, st);
  if (p->fused) return s2s_process_fused(p, d_traces, nfiles, rows, (float2 *)d_out, st);
  // K1 prefilter
  const int Mf = c.nfft_filter / 2;
  size_t lds = lds_bytes(Mf + 1, Mf);
  FXC_TRY(set_lds(s2s_prefilter_kernel, lds));
  const double inv_f = 1.0 / (double)((float)c.nfft_filter * c.delta);
  hipLaunchKernelGGL(s2s_prefilter_kernel, dim3((unsigned)rows), dim3(kS2sThreads), lds, st,
                     d_traces, c.npts, nfiles, p->d_steps, c.shift, c.seg, p->taper_size, p->tf,
                     p->d_resp, c.delta, inv_f, p->d_x);
  FXC_HIP(hipGetLastError());
  int mode = kLoadPlain;
  const float *w = nullptr;
  if (c.whiten == 1 || c.whiten == 3) FXC_TRY(s2s_whiten(p, rows, frames, kLoadPlain, nullptr, nullptr, st));
  if (c.normalize == 3) {
    FXC_TRY(launch_weight(false, p->d_x, c.seg, c.seg, c.num_ch, p->runabs_win, frames, p->d_w,
                          c.seg, st));
    mode = kLoadRunabs;
    w = p->d_w;
  } else if (c.normalize == 2) {
    mode = kLoadOnebit;
  } else if (c.normalize == 1) {
    const int nb = c.nfft_filter / 2 + 1;
    lds = lds_bytes(Mf + 1, Mf);
    FXC_TRY(set_lds(s2s_mf_base_kernel, lds));
    hipLaunchKernelGGL(s2s_mf_base_kernel, dim3((unsigned)rows), dim3(kS2sThreads), lds, st,
                       p->d_x, c.seg, p->tf, c.delta, p->d_base, p->bpitch);
    FXC_HIP(hipGetLastError());
    FXC_HIP(hipMemsetAsync(p->d_total, 0, (size_t)rows * c.seg * sizeof(float), st));
    FXC_TRY(set_lds(s2s_mf_band_kernel, lds));
    for (int i = 1; i < c.filter_count; ++i) {
      hipLaunchKernelGGL(s2s_mf_band_kernel, dim3((unsigned)rows), dim3(kS2sThreads), lds, st,
                         p->d_base, p->bpitch, p->tf, p->d_resp + (size_t)i * nb, inv_f, c.seg,
                         p->d_xb);
      FXC_HIP(hipGetLastError());
      FXC_TRY(launch_weight(false, p->d_xb, c.seg, c.seg, c.num_ch, p->mf_win[i], frames, p->d_w,
                            c.seg, st));
      const int nbr = (int)ceil_div(c.seg, 256);
      hipLaunchKernelGGL(s2s_mf_accum_kernel, dim3((unsigned)(rows * nbr)), dim3(256), 0, st,
                         p->d_xb, p->d_w, c.seg, c.num_ch, nbr, p->d_total);
      FXC_HIP(hipGetLastError());
    }
    FXC_HIP(hipMemcpyAsync(p->d_x, p->d_total, (size_t)rows * c.seg * sizeof(float),
                           hipMemcpyDeviceToDevice, st));
  }
  if (c.whiten == 2 || c.whiten == 3) {
    FXC_TRY(s2s_whiten(p, rows, frames, mode, w, (float2 *)d_out, st));
  } else {
    const int Mo = c.nfft_out / 2;
    lds = lds_bytes(Mo + 1, Mo);
    FXC_TRY(set_lds(s2s_output_kernel, lds));
    hipLaunchKernelGGL(s2s_output_kernel, dim3((unsigned)rows), dim3(kS2sThreads), lds, st,
                       p->d_x, c.seg, c.num_ch, mode, w, p->to, c.delta, c.phase_only,
                       (float2 *)d_out, p->nspec);
    FXC_HIP(hipGetLastError());
  }
  return FXC_OK;
}
