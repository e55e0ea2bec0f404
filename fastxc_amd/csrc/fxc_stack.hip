// Phase-weighted stacks: ncf_pws and ncf_tfpws hot paths.
//
// PWS replaces compute_pws_host_batch (native/pws/worker/gpu.cu:87-282):
// R2C + Hilbert mask -> C2C inverse -> a/N / max(|a/N|, 1e-7) -> W = sum_g
// w_g phi_g -> pws = lin * |W| (kernels/cuda.pws.cu:3-95, power 1).
// TF-PWS replaces compute_tfpws_from_prestack
// (native/tfpws/core/tfpws_compute.cu:19-366): for every voice f of the
// Stockwell transform the Gaussian-modulated, shifted analytic spectra of
// the linear stack and of every group are inverse transformed, the groups'
// unit phasors weight the linear stack's voice (|W|^1, the weight carrying
// 1/N), the voice is summed over time, optionally blended with the linear
// spectrum in a band, and the voices are inverse transformed
// (kernels/cuda.stransform.cu:59-332).
//
// MI355X design: the correlation length (2001 = 3*23*29) is transformed in
// LDS by 256-thread workgroups with register butterflies up to radix 29;
// lengths with other prime factors (2H+1 is arbitrary) use Bluestein's
// convolution of a factorable length m >= 2n-1, also in LDS, so every
// length up to 8192 (PWS) / 6144 (TF-PWS) is served.
// The analytic spectra of all groups are computed once; the TF-PWS voice
// kernel keeps the modulated transforms, the weight W_f(tau) and the
// Gaussian in LDS, so per voice only the (group-count + 1) spectra are read
// (from L2) and one complex number is written: the S-transform is never
// materialised in HBM.  W is accumulated in the FFT's digit-reversed order,
// which is a consistent permutation of tau, so no reordering pass exists.
#include "fxc_plan_cache.hpp"
#include "fastxc_amd.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace fxc {

constexpr int kStkThreads = 256;
constexpr int kStkBatch = 4;       // transforms held in LDS at once (at most)
constexpr int kVoiceChunk = 32;    // TF-PWS voices per workgroup
constexpr int kLdsReserve = 16;    // float2 slots left for static LDS
constexpr int kVoiceThreads = 512; // PWS and TF-PWS voice workgroups

// Length-n transform of the stacks: the direct mixed-radix plan when n
// factors into the implemented radices (rows of n values, digit-reversed
// output), else Bluestein's convolution of length m (rows of m values,
// natural-order output).
struct StackFftDev {
  FftPlanDev fwd;              // length ld, sign -1
  FftPlanDev inv;              // length ld, sign +1
  int n, ld, blue;
  const float2 *chirp;         // [n] exp(-i pi j^2 / n) (Bluestein only)
  const float2 *bhat_f;        // [m] forward-sign convolution kernel spectrum
  const float2 *bhat_i;        // [m] inverse-sign
};

struct StackFft {
  StackFftDev dev{};
  const float2 *tw_f = nullptr, *tw_i = nullptr;
  int twc = 0;                 // LDS twiddle slots of both tables
};

__device__ __forceinline__ int stack_pos(const StackFftDev &F, int k) {
  return F.blue ? k : fft_pos(F.fwd, k);
}

// twl: [fwd table | inv table]
__device__ __forceinline__ void stage_stack_tw(float2 *twl, const StackFftDev &F,
                                               const float2 *__restrict__ tw_f,
                                               const float2 *__restrict__ tw_i) {
  const int cf = kTwFine + F.fwd.ncoarse, ci = kTwFine + F.inv.ncoarse;
  for (int i = threadIdx.x; i < cf + ci; i += blockDim.x) twl[i] = i < cf ? tw_f[i] : tw_i[i - cf];
}

// In-place transforms (sgn -1 forward, +1 inverse) of nt rows at stride
// F.ld whose input is natural order in [0, n).  The caller synchronises
// before the call; the call ends with a barrier.  Output position of bin k:
// stack_pos(F, k).
__device__ __forceinline__ void stack_fft(float2 *buf, int nt, const StackFftDev &F, const float2 *twl, float sgn) {
  const float2 *twf = twl;
  const float2 *twi = twl + kTwFine + F.fwd.ncoarse;
  if (!F.blue) {
    if (sgn < 0.f)
      lds_fft<kRadixAll>(buf, nt, F.fwd, twf, -1.f);
    else
      lds_fft<kRadixAll>(buf, nt, F.inv, twi, 1.f);
    return;
  }
  const int n = F.n, m = F.ld, total = nt * m;
  const float cs = sgn < 0.f ? 1.f : -1.f;  // c_s = c_{-1} or its conjugate
  const float2 *bh = sgn < 0.f ? F.bhat_f : F.bhat_i;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int j = i % m;
    float2 v = make_float2(0.f, 0.f);
    if (j < n) {
      const float2 c = F.chirp[j];
      v = cmul(buf[i], make_float2(c.x, cs * c.y));
    }
    buf[i] = v;
  }
  __syncthreads();
  lds_fft<kRadixAll>(buf, nt, F.fwd, twf, -1.f);
  for (int i = threadIdx.x; i < total; i += blockDim.x) buf[i] = cmul(buf[i], bh[i % m]);
  __syncthreads();
  lds_fft_dit<kRadixAll>(buf, nt, F.inv, twi, 1.f);
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int j = i % m;
    if (j < n) {
      const float2 c = F.chirp[j];
      buf[i] = cmul(buf[i], make_float2(c.x, cs * c.y));
    }
  }
  __syncthreads();
}

// stack_fft with the direction and the Bluestein choice fixed at compile
// time.  In a kernel that inlines both directions of the runtime form the
// compiler copies the plan structure into scratch memory (976 B per lane at
// the plan sizes here); with both choices fixed, the plan is indexed in the
// kernel arguments.
template <int kSgn, bool kBlue>
__device__ __forceinline__ void stack_fft_dir(float2 *buf, int nt, const StackFftDev &F,
                                              const float2 *twl) {
  if constexpr (kBlue)
    stack_fft(buf, nt, F, twl, (float)kSgn);
  else if constexpr (kSgn < 0)
    lds_fft<kRadixAll>(buf, nt, F.fwd, twl, -1.f);
  else
    lds_fft<kRadixAll>(buf, nt, F.inv, twl + kTwFine + F.fwd.ncoarse, 1.f);
}

// Analytic spectra of real rows: A[k] = 2 X[k] for 1 <= k < (n+1)/2, else 0
// (hilbertTransformKernel after R2C).  Output natural order, [rows][n].
// Two real rows share one complex transform, z = s0 x0 + i s1 x1, whose
// spectrum separates as s0 A0[k] = Z[k] + conj Z[n-k] and
// s1 A1[k] = -i (Z[k] - conj Z[n-k]).  s0, s1 are powers of two that bring
// each row's largest |x| into [1, 2), so both rows carry the transform's
// rounding relative to their own size and the unscaling is exact.  per_wg
// transforms (2 per_wg rows) per workgroup.  A workgroup whose rows hold a
// non-finite value transforms its rows one by one instead, so such a value
// stays in its own row, as in a transform per row.
constexpr int kAnalyticMaxPer = 3;
constexpr int kLoadBatch = 8;  // loads a thread issues before it uses the first

// Row t[u] and column j[u] of flat index i0 + u step in rows of n, u < kB,
// from one division (a division per element costs a kernel that issues VALU
// most of its time several percent).
template <int kB>
__device__ __forceinline__ void flat_rows(int i0, int step, int n, int (&t)[kB], int (&j)[kB]) {
  int tt = i0 / n, jj = i0 - tt * n;
#pragma unroll
  for (int u = 0; u < kB; ++u) {
    t[u] = tt;
    j[u] = jj;
    jj += step;
    while (jj >= n) {
      jj -= n;
      ++tt;
    }
  }
}

// Loads of rows a kernel reads exactly once: nontemporal, so the twiddles
// and tables every workgroup reads stay in L2.
__device__ __forceinline__ float stream_load(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float2 stream_load(const float2 *p) {
  return __builtin_bit_cast(float2, __builtin_nontemporal_load((const unsigned long long *)p));
}

template <bool kBlue>
__global__ void __launch_bounds__(kStkThreads)
analytic_kernel(const float *__restrict__ x, long long rows, int per_wg, StackFftDev F,
                const float2 *__restrict__ tw_f, const float2 *__restrict__ tw_i,
                float2 *__restrict__ A) {
  extern __shared__ float2 lds[];
  __shared__ unsigned rmax[2 * kAnalyticMaxPer];  // float bits of max |x| per row
  __shared__ int bad;                             // a non-finite value in the rows
  const int n = F.n, ld = F.ld, tid = threadIdx.x;
  float2 *twl = lds + (size_t)per_wg * ld;
  stage_stack_tw(twl, F, tw_f, tw_i);
  const long long r0 = (long long)blockIdx.x * 2 * per_wg;
  const int nr = (int)std::min<long long>(2 * per_wg, rows - r0);
  const int nt = (nr + 1) / 2;
  if (tid < 2 * kAnalyticMaxPer) rmax[tid] = 0u;
  if (tid == 0) bad = 0;
  __syncthreads();
  bool nonfinite = false;
  float m[2 * kAnalyticMaxPer];
#pragma unroll
  for (int r = 0; r < 2 * kAnalyticMaxPer; ++r) m[r] = 0.f;
  // kLoadBatch positions' loads are issued before the first is used
  const int total = nt * n;
  for (int i0 = tid; i0 < total; i0 += kLoadBatch * (int)blockDim.x) {
    float av[kLoadBatch], bv[kLoadBatch];
    int tu[kLoadBatch], ju[kLoadBatch];
    flat_rows<kLoadBatch>(i0, (int)blockDim.x, n, tu, ju);
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      av[u] = bv[u] = 0.f;
      if (i < total) {
        const int t = tu[u], j = ju[u];
        av[u] = stream_load(x + (r0 + 2 * t) * n + j);
        if (2 * t + 1 < nr) bv[u] = stream_load(x + (r0 + 2 * t + 1) * n + j);
      }
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < total) {
        const int t = tu[u], j = ju[u];
        const float a = av[u], b = bv[u];
        lds[t * ld + j] = make_float2(a, b);
        nonfinite |= !isfinite(a) || !isfinite(b);
#pragma unroll
        for (int r = 0; r < kAnalyticMaxPer; ++r)
          if (r == t) {
            m[2 * r] = fmaxf(m[2 * r], fabsf(a));
            m[2 * r + 1] = fmaxf(m[2 * r + 1], fabsf(b));
          }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 2 * kAnalyticMaxPer; ++r)
    if (m[r] > 0.f) atomicMax(&rmax[r], __float_as_uint(m[r]));  // |x| >= 0: bits order as values
  if (nonfinite) bad = 1;
  __syncthreads();
  const int half = (n + 1) / 2;
  if (bad) {
    for (int h = 0; h < 2; ++h) {
      for (int t = 0; t < nt; ++t)
        for (int j = tid; j < n; j += blockDim.x) {
          const int r = 2 * t + h;
          lds[t * ld + j] = make_float2(r < nr ? x[(r0 + r) * n + j] : 0.f, 0.f);
        }
      __syncthreads();
      stack_fft_dir<-1, kBlue>(lds, nt, F, twl);
      for (int t = 0; t < nt; ++t)
      for (int k = tid; k < n; k += blockDim.x) {
        const int r = 2 * t + h;
        float2 v = make_float2(0.f, 0.f);
        if (k >= 1 && k < half) {
          const float2 z = lds[t * ld + stack_pos(F, k)];
          v = make_float2(2.f * z.x, 2.f * z.y);
        }
        if (r < nr) A[(r0 + r) * n + k] = v;
      }
      __syncthreads();
    }
    return;
  }
  // scales 2^-e with max |x| in [2^(e-1), 2^e); 1 for an all-zero row (or
  // a non-finite maximum, which then propagates as in a single transform)
  float sc[2 * kAnalyticMaxPer];
#pragma unroll
  for (int r = 0; r < 2 * kAnalyticMaxPer; ++r) {
    const float mx = __uint_as_float(rmax[r]);
    int e = 0;
    if (mx > 0.f && isfinite(mx)) (void)frexpf(mx, &e);
    sc[r] = ldexpf(1.f, -e);
  }
#pragma unroll
  for (int t = 0; t < kAnalyticMaxPer; ++t) {
    if (t >= nt) break;
    const float s0 = sc[2 * t], s1 = sc[2 * t + 1];
    for (int j = tid; j < n; j += blockDim.x) {
      const float2 v = lds[t * ld + j];
      lds[t * ld + j] = make_float2(v.x * s0, v.y * s1);
    }
  }
  __syncthreads();
  stack_fft_dir<-1, kBlue>(lds, nt, F, twl);
#pragma unroll
  for (int t = 0; t < kAnalyticMaxPer; ++t) {
    if (t >= nt) break;
    const float u0 = 1.f / sc[2 * t], u1 = 1.f / sc[2 * t + 1];
    for (int k = tid; k < n; k += blockDim.x) {
      float2 a0 = make_float2(0.f, 0.f), a1 = make_float2(0.f, 0.f);
      if (k >= 1 && k < half) {
        const float2 zk = lds[t * ld + stack_pos(F, k)];
        const float2 zm = lds[t * ld + stack_pos(F, n - k)];
        a0 = make_float2(zk.x + zm.x, zk.y - zm.y);              // Z[k] + conj Z[n-k]
        const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);  // Z[k] - conj Z[n-k]
        a1 = make_float2(d.y, -d.x);                             // -i d
      }
      A[(r0 + 2 * t) * n + k] = make_float2(a0.x * u0, a0.y * u0);
      if (2 * t + 1 < nr) A[(r0 + 2 * t + 1) * n + k] = make_float2(a1.x * u1, a1.y * u1);
    }
  }
}

// |W|^power from |W|^2: power 1 (the reference) is the square root, power 2
// the square itself, other powers powf (spec mode).
__device__ __forceinline__ float mag_pow(float mag2, float power) {
  if (power == 1.f) return sqrtf(mag2);
  if (power == 2.f) return mag2;
  return mag2 > 0.f ? powf(mag2, 0.5f * power) : 0.f;
}

// PWS: one workgroup per pair.  Thread t owns transform positions
// p = t + k kThreads (k < kPer), so W stays in registers and the LDS holds
// only the group rows and the twiddles.
template <int kThreads, int kPer>
__global__ void __launch_bounds__(kThreads)
pws_kernel(const float2 *__restrict__ A, const float *__restrict__ weights,
           const long long *__restrict__ offsets, const float *__restrict__ lin, StackFftDev F,
           const float2 *__restrict__ tw_f, const float2 *__restrict__ tw_i, int nbatch,
           float power, float *__restrict__ out) {
  extern __shared__ float2 lds[];
  const int n = F.n, ld = F.ld;
  float2 *buf = lds;                     // [nbatch][ld]
  float2 *twl = buf + (size_t)nbatch * ld;
  stage_stack_tw(twl, F, tw_f, tw_i);
  const int tid = threadIdx.x;
  float2 W[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) W[k] = make_float2(0.f, 0.f);
  const int pair = blockIdx.x;
  const long long g0 = offsets[pair], g1 = offsets[pair + 1];
  const float invn = 1.f / (float)n;
  for (long long gb = g0; gb < g1; gb += nbatch) {
    const int nb = (int)std::min<long long>(nbatch, g1 - gb);
    __syncthreads();
    // half a batch: W already holds 2 kPer registers here
    constexpr int kB = kLoadBatch / 2;
    for (int i0 = tid; i0 < nb * n; i0 += kB * kThreads) {
      float2 v[kB];
      int tu[kB], ju[kB];
      flat_rows<kB>(i0, kThreads, n, tu, ju);
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int i = i0 + u * kThreads;
        if (i < nb * n) v[u] = stream_load(A + (gb + tu[u]) * n + ju[u]);
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int i = i0 + u * kThreads;
        if (i < nb * n) buf[tu[u] * ld + ju[u]] = v[u];
      }
    }
    __syncthreads();
    stack_fft(buf, nb, F, twl, 1.f);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int p = tid + k * kThreads;
      if (p < n) {
        float2 acc = W[k];
        for (int b = 0; b < nb; ++b) {
          const float2 z = buf[b * ld + p];
          const float re = z.x * invn, im = z.y * invn;
          float mod = sqrtf(re * re + im * im);
          mod = mod > 1e-7f ? mod : 1e-7f;
          const float wg = weights[gb + b];
          acc.x += (re / mod) * wg;
          acc.y += (im / mod) * wg;
        }
        W[k] = acc;
      }
    }
  }
  __syncthreads();
  // W is indexed by transform position: publish it in LDS, read it by lag
  float2 *Ws = buf;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int p = tid + k * kThreads;
    if (p < n) Ws[p] = W[k];
  }
  __syncthreads();
  const float *l = lin + (size_t)pair * n;
  float *o = out + (size_t)pair * n;
  for (int t = tid; t < n; t += kThreads) {
    const float2 w = Ws[stack_pos(F, t)];
    o[t] = l[t] * mag_pow(w.x * w.x + w.y * w.y, power);
  }
}

// TF-PWS voices [f0, f0 + kVoiceChunk) of one pair: out_spec[pair][f] =
// sum_tau ST_lin,f(tau) |W_f(tau)|.  The LDS holds `nrows` rows; groups go
// through them `nbatch` at a time, and the linear stack's modulated
// spectrum rides in the row after the last batch's groups when one is free
// (then a voice with few groups is a single transform call over G + 1 rows,
// whose large-radix passes have enough butterflies to occupy the threads),
// else it is transformed alone after them.  Thread t owns transform
// positions p = t + k kThreads (k < kPer), so W_f stays in its registers;
// LDS holds only the Gaussian, the rows and the twiddles (74 KB at n = 2001
// with three groups: two workgroups per CU).
template <int kThreads, int kPer>
__global__ void __launch_bounds__(kThreads)
tfpws_voice_kernel(const float2 *__restrict__ Ag, const float2 *__restrict__ Alin,
                   const float *__restrict__ weights, const long long *__restrict__ offsets,
                   StackFftDev F, const float2 *__restrict__ tw_f, const float2 *__restrict__ tw_i,
                   int nrows, int nbatch, int nfreq, float power,
                   const unsigned char *__restrict__ chunk_active,
                   const float *__restrict__ gtab, float2 *__restrict__ out_spec) {
  extern __shared__ float2 lds[];
  const int n = F.n, ld = F.ld;
  float *gauss = (float *)lds;                       // [n] floats (n/2 complex slots)
  float2 *buf = lds + (n + 1) / 2;                   // [nrows][ld]
  float2 *twl = buf + (size_t)nrows * ld;
  __shared__ float red_x[kThreads / 64], red_y[kThreads / 64];
  const int pair = blockIdx.y;
  const int chunk = blockIdx.x;
  if (chunk_active && !chunk_active[chunk]) return;  // uniform per workgroup
  stage_stack_tw(twl, F, tw_f, tw_i);
  const int f0 = chunk * kVoiceChunk;
  const long long g0 = offsets[pair], g1 = offsets[pair + 1];
  const float2 *Al = Alin + (size_t)pair * n;
  const int half = (n + 1) / 2;
  const float nf = (float)n;
  const int tid = threadIdx.x;
  for (int f = f0; f < std::min(nfreq, f0 + kVoiceChunk); ++f) {
    __syncthreads();
    // Gaussian window of voice f (gaussianModulateSub, scale 0.1), from the
    // table of tfpws_gauss_kernel: g(c) = gtab[f][min(c, n - c)]
    const float *gf = gtab + (size_t)f * (half + 1);
    for (int c = tid; c < n; c += kThreads) gauss[c] = gf[c < half ? c : n - c];
    float2 W[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) W[k] = make_float2(0.f, 0.f);
    // batches of nbatch groups; the linear stack joins the last one when a
    // row is free there, else it forms one more batch on its own (one call
    // site for every transform keeps the kernel's registers at one copy)
    const long long G = g1 - g0;
    const long long nbat = (G + nbatch - 1) / nbatch;
    const bool lin_joins = G - (nbat - 1) * nbatch < nrows;
    const long long total = nbat + (lin_joins ? 0 : 1);
    int lin_row = 0;
    for (long long bi = 0; bi < total; ++bi) {
      const long long gb = g0 + bi * nbatch;
      const int nb = bi < nbat ? (int)std::min<long long>(nbatch, g1 - gb) : 0;
      const int rows = nb + (bi == total - 1 ? 1 : 0);  // + linear stack in the last
      __syncthreads();
      for (int row = 0; row < rows; ++row) {
        const float2 *arow = row == nb ? Al : Ag + (gb + row) * n;
        for (int c = tid; c < n; c += kThreads) {
          int src = f + c;
          src = src >= n ? src - n : src;
          buf[row * ld + c] = cscale(arow[src], gauss[c]);
        }
      }
      __syncthreads();
      stack_fft(buf, rows, F, twl, 1.f);
      lin_row = nb;
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int p = tid + k * kThreads;
        if (p < n) {
          float2 acc = W[k];
          for (int b = 0; b < nb; ++b) {
            const float2 z = buf[b * ld + p];
            float fa = sqrtf(z.x * z.x + z.y * z.y);
            if (fa <= 1e-7f) fa = 1.f;
            // w z / (|z| N) with one division per position
            const float r = weights[gb + b] / (fa * nf);
            acc.x += z.x * r;
            acc.y += z.y * r;
          }
          W[k] = acc;
        }
      }
    }
    const float2 *L = buf + (size_t)lin_row * ld;
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int p = tid + k * kThreads;
      if (p < n) {
        const float mag = mag_pow(W[k].x * W[k].x + W[k].y * W[k].y, power);
        sx += L[p].x * mag;
        sy += L[p].y * mag;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      sx += __shfl_down(sx, off);
      sy += __shfl_down(sy, off);
    }
    if ((tid & 63) == 0) {
      red_x[tid >> 6] = sx;
      red_y[tid >> 6] = sy;
    }
    __syncthreads();
    if (tid == 0) {
      float tx = 0.f, ty = 0.f;
      for (int i = 0; i < kThreads / 64; ++i) {
        tx += red_x[i];
        ty += red_y[i];
      }
      out_spec[(size_t)pair * n + f] = make_float2(tx, ty);
    }
  }
}

// Gaussian window of voice f (gaussianModulateSub, scale 0.1) into gauss[0, n).
__device__ __forceinline__ void tfpws_gaussian(float *gauss, int n, int f) {
  const int half = (n + 1) / 2;
  const float pi2 = -2.0f * (float)M_PI * (float)M_PI;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    float g;
    if (f == 0) {
      g = 1.f;
    } else if (c == 0) {
      g = 0.f;
    } else {
      const int cm = c < half ? c : n - c;
      g = cm < half ? expf(pi2 * (float)cm * (float)cm * 0.1f / ((float)f * (float)f)) : 0.f;
    }
    gauss[c] = g;
  }
}

// c[k] = |W_v(p)|^power at the positions p = tid + k kThreads this thread
// owns, W_v = sum_g w_g ST_g,v / (|ST_g,v|' n), the groups transformed up to
// nrows at a time.  Starts and ends with a barrier.
template <int kThreads, int kPer>
__device__ __forceinline__ void tfpws_voice_weight(float (&c)[kPer], int v, const float2 *Ag,
                                                   const float *weights, long long g0,
                                                   long long g1, const StackFftDev &F,
                                                   const float2 *twl, float *gauss, float2 *buf,
                                                   int nrows, float power) {
  const int n = F.n, ld = F.ld, tid = threadIdx.x;
  const float nf = (float)n;
  __syncthreads();
  tfpws_gaussian(gauss, n, v);
  float2 W[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) W[k] = make_float2(0.f, 0.f);
  for (long long gb = g0; gb < g1; gb += nrows) {
    const int nb = (int)std::min<long long>(nrows, g1 - gb);
    __syncthreads();
    for (int row = 0; row < nb; ++row)
      for (int cc = tid; cc < n; cc += kThreads) {
        int src = v + cc;
        src = src >= n ? src - n : src;
        buf[row * ld + cc] = cscale(Ag[(gb + row) * n + src], gauss[cc]);
      }
    __syncthreads();
    stack_fft(buf, nb, F, twl, 1.f);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int p = tid + k * kThreads;
      if (p < n) {
        float2 acc = W[k];
        for (int b = 0; b < nb; ++b) {
          const float2 z = buf[b * ld + p];
          float fa = sqrtf(z.x * z.x + z.y * z.y);
          if (fa <= 1e-7f) fa = 1.f;
          const float wg = weights[gb + b];
          acc.x += wg * z.x / (fa * nf);
          acc.y += wg * z.y / (fa * nf);
        }
        W[k] = acc;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) c[k] = mag_pow(W[k].x * W[k].x + W[k].y * W[k].y, power);
  __syncthreads();
}

// Spec-mode TF-PWS (no reference counterpart; the reference transforms the
// groups at every one of the n/2+1 voices, tfpws_compute.cu:224).  The
// groups are transformed only at nvoices voice bins
// v_j = floor(j (nfreq-1) / (nvoices-1)); every bin f still takes the linear
// stack's voice, weighted by c_f = (1-t) |W_vj|^p + t |W_vj+1|^p with
// t = (f - v_j) / (v_j+1 - v_j) and v_j <= f < v_j+1.  One workgroup per
// (pair, 32 consecutive bins) walks its bins in order and keeps the weights
// of the current pair of voices in registers, so each voice is transformed
// once per chunk that needs it.
template <int kThreads, int kPer>
__global__ void __launch_bounds__(kThreads)
tfpws_interp_kernel(const float2 *__restrict__ Ag, const float2 *__restrict__ Alin,
                    const float *__restrict__ weights, const long long *__restrict__ offsets,
                    StackFftDev F, const float2 *__restrict__ tw_f, const float2 *__restrict__ tw_i,
                    int nrows, int nfreq, int nvoices, float power,
                    const unsigned char *__restrict__ chunk_active, float2 *__restrict__ out_spec) {
  extern __shared__ float2 lds[];
  const int n = F.n, ld = F.ld;
  float *gauss = (float *)lds;                       // [n] floats
  float2 *buf = lds + (n + 1) / 2;                   // [nrows][ld]
  float2 *twl = buf + (size_t)nrows * ld;
  __shared__ float red_x[kThreads / 64], red_y[kThreads / 64];
  const int pair = blockIdx.y;
  const int chunk = blockIdx.x;
  if (chunk_active && !chunk_active[chunk]) return;  // uniform per workgroup
  stage_stack_tw(twl, F, tw_f, tw_i);
  const int f0 = chunk * kVoiceChunk;
  const int f1 = std::min(nfreq, f0 + kVoiceChunk);
  const long long g0 = offsets[pair], g1 = offsets[pair + 1];
  const float2 *Al = Alin + (size_t)pair * n;
  const int tid = threadIdx.x;
  const long long va = nfreq - 1, vb = nvoices - 1;
  float cL[kPer], cR[kPer];
  int jL = -1, jR = -1;  // voices whose weights cL / cR hold
  for (int f = f0; f < f1; ++f) {
    // largest j with v_j <= f: floor(j a / b) <= f  <=>  j <= ((f+1) b - 1) / a
    const int j = (int)std::min<long long>(vb, ((long long)(f + 1) * vb - 1) / va);
    const int vj = (int)((long long)j * va / vb);
    if (j != jL) {
      if (j == jR) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) cL[k] = cR[k];
      } else {
        tfpws_voice_weight<kThreads, kPer>(cL, vj, Ag, weights, g0, g1, F, twl, gauss, buf,
                                           nrows, power);
      }
      jL = j;
    }
    float t = 0.f;
    if (vj != f) {
      const int vj1 = (int)((long long)(j + 1) * va / vb);
      if (jR != j + 1) {
        tfpws_voice_weight<kThreads, kPer>(cR, vj1, Ag, weights, g0, g1, F, twl, gauss, buf,
                                           nrows, power);
        jR = j + 1;
      }
      t = (float)(f - vj) / (float)(vj1 - vj);
    }
    // the linear stack's voice f
    __syncthreads();
    tfpws_gaussian(gauss, n, f);
    __syncthreads();
    for (int cc = tid; cc < n; cc += kThreads) {
      int src = f + cc;
      src = src >= n ? src - n : src;
      buf[cc] = cscale(Al[src], gauss[cc]);
    }
    __syncthreads();
    stack_fft(buf, 1, F, twl, 1.f);
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int p = tid + k * kThreads;
      if (p < n) {
        const float c = t == 0.f ? cL[k] : (1.f - t) * cL[k] + t * cR[k];
        sx += buf[p].x * c;
        sy += buf[p].y * c;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      sx += __shfl_down(sx, off);
      sy += __shfl_down(sy, off);
    }
    if ((tid & 63) == 0) {
      red_x[tid >> 6] = sx;
      red_y[tid >> 6] = sy;
    }
    __syncthreads();
    if (tid == 0) {
      float tx = 0.f, ty = 0.f;
      for (int i = 0; i < kThreads / 64; ++i) {
        tx += red_x[i];
        ty += red_y[i];
      }
      out_spec[(size_t)pair * n + f] = make_float2(tx, ty);
    }
  }
}

// ---------------------------------------------------------------------------
// Running PWS state of the streaming engine (no reference counterpart: the
// reference stacks a pair's stored records offline in sub-stacks of B,
// native/pws/io/sourcepack_io.cu:385-478, kernels/cuda.pws.cu:28-95).  Per
// pair the engine keeps S[t] = sum over closed sub-stacks g of n_g phi_g[t]
// and the float sum of the open sub-stack's records; then
//   W = (S + n_open phi_open) / T,   pws = lin |W|^power
// equals the reference's sum_g (n_g / T) phi_g over all T records.
// phi of a sub-stack: its mean (sum * (1/n), float), analytic signal (bins
// 1..ceil(N/2)-1 doubled, the rest 0), unnormalised inverse / N, divided by
// max(|a|, 1e-7) -- the steps of analytic_kernel and pws_kernel.
// ---------------------------------------------------------------------------
template <int kPer>
__device__ __forceinline__ void stream_phi(float2 *lds, const float *__restrict__ sum, float scale,
                                           const StackFftDev &F, const float2 *twl,
                                           float2 (&phi)[kPer]) {
  const int n = F.n, tid = threadIdx.x;
  for (int j = tid; j < F.ld; j += blockDim.x)
    lds[j] = j < n ? make_float2(sum[j] * scale, 0.f) : make_float2(0.f, 0.f);
  __syncthreads();
  stack_fft(lds, 1, F, twl, -1.f);
  // analytic mask into natural order (through registers)
  const int half = (n + 1) / 2;
  float2 a[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int f = tid + k * kStkThreads;
    a[k] = make_float2(0.f, 0.f);
    if (f >= 1 && f < half) {
      const float2 z = lds[stack_pos(F, f)];
      a[k] = make_float2(2.f * z.x, 2.f * z.y);
    }
  }
  __syncthreads();
  for (int j = tid; j < F.ld; j += blockDim.x) lds[j] = make_float2(0.f, 0.f);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int f = tid + k * kStkThreads;
    if (f < n) lds[f] = a[k];
  }
  __syncthreads();
  stack_fft(lds, 1, F, twl, 1.f);
  const float invn = 1.f / (float)n;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int t = tid + k * kStkThreads;
    phi[k] = make_float2(0.f, 0.f);
    if (t < n) {
      const float2 z = lds[stack_pos(F, t)];
      const float re = z.x * invn, im = z.y * invn;
      float mod = sqrtf(re * re + im * im);
      mod = mod > 1e-7f ? mod : 1e-7f;
      phi[k] = make_float2(re / mod, im / mod);
    }
  }
  __syncthreads();
}

// Closes up to kCloseRows sub-stacks per workgroup (per_wg of them):
// S[slot] += cnt phi(open[slot] / cnt), open[slot] = 0, with phi the steps
// of stream_phi.  With rows (sub-stacks of one record): the t-th closing
// sub-stack is row t of rows itself, cnt is 1, and no open sum is kept.  At n = 2001 one row keeps only 69-87 of the 256 threads
// busy in the radix-23 and radix-29 passes, so the rows of a workgroup share
// each transform call (as in analytic_kernel) and the twiddles are staged
// once for all of them.
constexpr int kCloseRows = 3;

template <int kPer, bool kBlue>
__global__ void __launch_bounds__(kStkThreads)
stream_pws_close_kernel(float2 *__restrict__ S, float *__restrict__ open,
                        const int *__restrict__ slot, const int *__restrict__ cnt,
                        const float *__restrict__ rows, long long nclose, int per_wg,
                        StackFftDev F, const float2 *__restrict__ tw_f,
                        const float2 *__restrict__ tw_i) {
  extern __shared__ float2 lds[];
  const int n = F.n, ld = F.ld, tid = threadIdx.x;
  float2 *twl = lds + (size_t)per_wg * ld;
  stage_stack_tw(twl, F, tw_f, tw_i);
  const long long r0 = (long long)blockIdx.x * per_wg;
  const int nt = (int)std::min<long long>(per_wg, nclose - r0);
  for (int i0 = tid; i0 < nt * ld; i0 += kLoadBatch * kStkThreads) {
    float v[kLoadBatch];
    int tu[kLoadBatch], ju[kLoadBatch];
    flat_rows<kLoadBatch>(i0, kStkThreads, ld, tu, ju);
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int i = i0 + u * kStkThreads;
      const int t = tu[u], j = ju[u];
      v[u] = 0.f;
      if (i < nt * ld && j < n)
        v[u] = rows ? stream_load(rows + (r0 + t) * n + j)
                    : open[(long long)slot[r0 + t] * n + j] * (1.0f / (float)cnt[r0 + t]);
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int i = i0 + u * kStkThreads;
      if (i < nt * ld) lds[i] = make_float2(v[u], 0.f);
    }
  }
  __syncthreads();
  stack_fft_dir<-1, kBlue>(lds, nt, F, twl);
  // analytic mask into natural order (through registers)
  const int half = (n + 1) / 2;
  float2 a[kCloseRows][kPer];
#pragma unroll
  for (int t = 0; t < kCloseRows; ++t)
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int f = tid + k * kStkThreads;
      a[t][k] = make_float2(0.f, 0.f);
      if (t < nt && f >= 1 && f < half) {
        const float2 z = lds[t * ld + stack_pos(F, f)];
        a[t][k] = make_float2(2.f * z.x, 2.f * z.y);
      }
    }
  __syncthreads();
  for (int i = tid; i < nt * ld; i += kStkThreads) lds[i] = make_float2(0.f, 0.f);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kCloseRows; ++t)
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int f = tid + k * kStkThreads;
      if (t < nt && f < n) lds[t * ld + f] = a[t][k];
    }
  __syncthreads();
  stack_fft_dir<1, kBlue>(lds, nt, F, twl);
  const float invn = 1.f / (float)n;
  for (int t = 0; t < nt; ++t) {
    const long long sl = slot[r0 + t];
    const float fc = rows ? 1.f : (float)cnt[r0 + t];
    float2 *Sr = S + sl * n;
    float *o = rows ? nullptr : open + sl * n;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int j = tid + k * kStkThreads;
      if (j < n) {
        const float2 z = lds[t * ld + stack_pos(F, j)];
        const float re = z.x * invn, im = z.y * invn;
        float mod = sqrtf(re * re + im * im);
        mod = mod > 1e-7f ? mod : 1e-7f;
        const float2 v = Sr[j];
        Sr[j] = make_float2(v.x + fc * (re / mod), v.y + fc * (im / mod));
        if (o) o[j] = 0.f;
      }
    }
  }
}

// Read-out per pair (one workgroup each): out = (float)(acc / T) |W|^power,
// W = (S + n_open phi_open) / T.
template <int kPer>
__global__ void __launch_bounds__(kStkThreads)
stream_pws_read_kernel(const float2 *__restrict__ S, const float *__restrict__ open,
                       const double *__restrict__ acc, const int *__restrict__ slot,
                       const long long *__restrict__ total, const int *__restrict__ nopen,
                       float power, StackFftDev F, const float2 *__restrict__ tw_f,
                       const float2 *__restrict__ tw_i, float *__restrict__ out) {
  extern __shared__ float2 lds[];
  const int n = F.n;
  float2 *twl = lds + F.ld;
  stage_stack_tw(twl, F, tw_f, tw_i);
  const int pair = blockIdx.x;
  const long long sl = slot[pair];
  const long long T = total[pair];
  const int no = nopen[pair];
  float2 phi[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) phi[k] = make_float2(0.f, 0.f);
  if (no > 0) stream_phi<kPer>(lds, open + sl * n, 1.0f / (float)no, F, twl, phi);
  const float invT = 1.0f / (float)T;
  const float fo = (float)no;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int t = threadIdx.x + k * kStkThreads;
    if (t < n) {
      const float2 sv = S[sl * n + t];
      const float wx = (sv.x + fo * phi[k].x) * invT, wy = (sv.y + fo * phi[k].y) * invT;
      const float lin = (float)(acc[sl * n + t] / (double)T);
      out[(size_t)pair * n + t] = lin * mag_pow(wx * wx + wy * wy, power);
    }
  }
}

// Final TF-PWS step per pair: voices (blended with the linear analytic
// spectrum in band mode), zero above nfreq, inverse transform, Re / n.
__global__ void __launch_bounds__(kStkThreads)
tfpws_final_kernel(const float2 *__restrict__ out_spec, const float2 *__restrict__ Alin, int nfreq,
                   int band_limited, float df, float fmin, float fmax, float taper, StackFftDev F,
                   const float2 *__restrict__ tw_f, const float2 *__restrict__ tw_i,
                   float *__restrict__ y) {
  extern __shared__ float2 lds[];
  const int n = F.n;
  float2 *twl = lds + F.ld;
  stage_stack_tw(twl, F, tw_f, tw_i);
  const int pair = blockIdx.x;
  const float2 *os = out_spec + (size_t)pair * n;
  const float2 *Al = Alin + (size_t)pair * n;
  for (int f = threadIdx.x; f < n; f += blockDim.x) {
    float2 v = make_float2(0.f, 0.f);
    if (band_limited) {
      v = Al[f];
      if (f < nfreq) {
        const float freq = f * df;
        float alpha = 0.f;
        if (freq >= fmin && freq <= fmax) {
          const float t = fminf(taper, 0.5f * (fmax - fmin));
          if (t <= 0.f)
            alpha = 1.f;
          else if (freq < fmin + t)
            alpha = 0.5f - 0.5f * cosf((float)M_PI * ((freq - fmin) / t));
          else if (freq > fmax - t)
            alpha = 0.5f - 0.5f * cosf((float)M_PI * ((fmax - freq) / t));
          else
            alpha = 1.f;
        }
        v = make_float2(Al[f].x * (1.f - alpha) + os[f].x * alpha,
                        Al[f].y * (1.f - alpha) + os[f].y * alpha);
      }
    } else if (f < nfreq) {
      v = os[f];
    }
    lds[f] = v;
  }
  __syncthreads();
  stack_fft(lds, 1, F, twl, 1.f);
  float *o = y + (size_t)pair * n;
  for (int t = threadIdx.x; t < n; t += blockDim.x) o[t] = lds[stack_pos(F, t)].x / (float)n;
}

// Linear stack of every pair: out[p][j] = (float)((sum_r (double)x[r][j]) / count_p)
// over rows r = off[p] .. off[p+1]-1 in row order (a float64 running sum,
// then a float64 division and one rounding, as _linear_stack_group does).
// One thread per sample; a wave reads 64 consecutive samples of a row.
constexpr int kLinThreads = 256;

__global__ void __launch_bounds__(kLinThreads)
linear_stack_kernel(const float *__restrict__ x, const long long *__restrict__ off, int n,
                    float *__restrict__ out) {
  const int j = blockIdx.x * kLinThreads + threadIdx.x;
  const int pair = blockIdx.y;
  if (j >= n) return;
  const long long r0 = off[pair], r1 = off[pair + 1];
  double acc = 0.0;
  const float *p = x + (size_t)r0 * n + j;
  long long r = r0;
  // eight records' loads in flight at once, added in record order
  for (; r + 8 <= r1; r += 8, p += 8 * (size_t)n) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = p[i * (size_t)n];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc += (double)v[i];
  }
  for (; r < r1; ++r, p += n) acc += (double)*p;
  out[(size_t)pair * n + j] = (float)(acc / (double)(r1 - r0));
}

struct StackWork {
  float2 *A = nullptr;      // [groups + pairs][n]
  long long *offsets = nullptr;
  float2 *spec = nullptr;   // [pairs][n]
  unsigned char *active = nullptr;
  float *gauss = nullptr;   // [nfreq][(n+1)/2 + 1] Gaussian windows of the voices
  ~StackWork() {
    dfree(A);
    dfree(offsets);
    dfree(spec);
    dfree(active);
    dfree(gauss);
  }
};

// Direct plan when n factors, else the Bluestein plan.  create = false
// only looks the plans up (no device allocation).
static int stack_fft_plan(int n, StackFft *S, bool create = true) {
  int radix[kMaxPasses], np = 0;
  const FftPlan *pf = nullptr, *pi = nullptr;
  StackFftDev &d = S->dev;
  if (factor_radices(n, radix, &np)) {
    FXC_TRY(plan_cache_complex(n, -1, &pf, 32, create));
    FXC_TRY(plan_cache_complex(n, 1, &pi, 32, create));
    d.blue = 0;
    d.ld = n;
    d.chirp = d.bhat_f = d.bhat_i = nullptr;
  } else {
    const BluesteinPlan *bp = nullptr;
    FXC_TRY(plan_cache_bluestein(n, &bp, create));
    pf = bp->fwd;
    pi = bp->inv;
    d.blue = 1;
    d.ld = bp->m;
    d.chirp = bp->d_chirp;
    d.bhat_f = bp->d_bhat_f;
    d.bhat_i = bp->d_bhat_i;
  }
  d.n = n;
  d.fwd = pf->dev;
  d.inv = pi->dev;
  S->tw_f = pf->d_tw;
  S->tw_i = pi->d_tw;
  S->twc = 2 * tw_lds_count(d.ld);
  return FXC_OK;
}

// Rows of `ld` the LDS holds beside `fixed` other slots and `extra_rows`
// reserved rows, capped by the largest group count of the call (0 when none
// fit).
static int stack_batch(const StackFft &S, size_t fixed, int extra_rows, const long long *off,
                       int npairs) {
  long long gmax = 1;
  for (int i = 0; i < npairs; ++i) gmax = std::max(gmax, off[i + 1] - off[i]);
  const long long room = (long long)kLdsComplexCapacity - kLdsReserve - (long long)fixed - S.twc -
                         (long long)extra_rows * S.dev.ld;
  if (room < S.dev.ld) return 0;
  return (int)std::min<long long>(std::min<long long>(kStkBatch, gmax), room / S.dev.ld);
}

// Device offsets and the analytic spectra of every group and (with_lin,
// TF-PWS) of every linear stack, into A [groups + pairs][n] and offsets
// [npairs + 1] (caller's buffers).  PWS weights only the groups' phases, so
// it leaves the linear stacks' rows out.
static int stack_analytic(const float *d_pre, const float *d_lin, const long long *h_off,
                          int npairs, int n, float2 *A, long long *offsets, const StackFft &S,
                          hipStream_t st, bool with_lin) {
  const long long ng = h_off[npairs];
  FXC_HIP(hipMemcpyAsync(offsets, h_off, (size_t)(npairs + 1) * sizeof(long long),
                         hipMemcpyHostToDevice, st));
  // at most 3 transforms (6 rows) per workgroup: three workgroups share a CU
  // at n = 2001
  const int per = std::max(1, std::min(kAnalyticMaxPer, (kLdsComplexCapacity - kLdsReserve - S.twc) /
                                                            S.dev.ld));
  const size_t lds = ((size_t)per * S.dev.ld + S.twc) * sizeof(float2);
  auto *kern = S.dev.blue ? analytic_kernel<true> : analytic_kernel<false>;
  FXC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  if (ng > 0) {
    hipLaunchKernelGGL(kern, dim3((unsigned)ceil_div((size_t)ng, (size_t)(2 * per))),
                       dim3(kStkThreads), lds, st, d_pre, ng, per, S.dev, S.tw_f, S.tw_i, A);
    FXC_HIP(hipGetLastError());
  }
  if (!with_lin) return FXC_OK;
  hipLaunchKernelGGL(kern, dim3((unsigned)ceil_div((size_t)npairs, (size_t)(2 * per))),
                     dim3(kStkThreads), lds, st, d_lin, (long long)npairs, per, S.dev, S.tw_f,
                     S.tw_i, A + (size_t)ng * n);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}

static int stack_common(const float *d_pre, const float *d_lin, const long long *h_off, int npairs,
                        int n, StackWork &w, const StackFft &S, hipStream_t st, bool with_lin) {
  const long long rows = h_off[npairs] + (with_lin ? npairs : 0);
  FXC_TRY(dmalloc((void **)&w.A, (size_t)std::max<long long>(rows, 1) * n * sizeof(float2)));
  FXC_TRY(dmalloc((void **)&w.offsets, (size_t)(npairs + 1) * sizeof(long long)));
  return stack_analytic(d_pre, d_lin, h_off, npairs, n, w.A, w.offsets, S, st, with_lin);
}

// The voice kernel's per-thread W registers come in steps of 4 positions.
// Gaussian windows of every voice for tfpws_voice_kernel, by the shift's
// distance cm = min(c, n - c) from bin 0: gtab[f][cm], cm in [0, (n+1)/2],
// the values the voice kernel would form (gaussianModulateSub, scale 0.1):
// 1 at voice 0, 0 at c = 0, and 0 at cm = n/2 for even n, where the shift
// meets neither branch of the reference (c < half, or a mirror n - c < half).
__global__ void __launch_bounds__(kStkThreads)
tfpws_gauss_kernel(int n, int nfreq, float *__restrict__ gtab) {
  const int half = (n + 1) / 2;
  const float pi2 = -2.0f * (float)M_PI * (float)M_PI;
  const int f = blockIdx.y;
  for (int cm = blockIdx.x * kStkThreads + threadIdx.x; cm <= half; cm += gridDim.x * kStkThreads) {
    float g;
    if (f == 0)
      g = 1.f;
    else if (cm == 0)
      g = 0.f;
    else
      g = cm < half ? expf(pi2 * (float)cm * (float)cm * 0.1f / ((float)f * (float)f)) : 0.f;
    gtab[(size_t)f * (half + 1) + cm] = g;
  }
}

static int launch_voice(int n, size_t lds, dim3 grid, hipStream_t st, const float2 *Ag,
                        const float2 *Alin, const float *wts, const long long *off,
                        const StackFft &S, int nrows, int nbatch, int nfreq, float power,
                        int nvoices, const unsigned char *active, const float *gtab,
                        float2 *spec) {
  const int per = (n + kVoiceThreads - 1) / kVoiceThreads;
#define FXC_VOICE(P)                                                                        \
  if (per <= P) {                                                                          \
    if (nvoices > 0) {                                                                     \
      auto *k = tfpws_interp_kernel<kVoiceThreads, P>;                                     \
      FXC_HIP(hipFuncSetAttribute((const void *)k,                                         \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));  \
      hipLaunchKernelGGL(k, grid, dim3(kVoiceThreads), lds, st, Ag, Alin, wts, off, S.dev, \
                         S.tw_f, S.tw_i, nbatch, nfreq, nvoices, power, active, spec);     \
    } else {                                                                               \
      auto *k = tfpws_voice_kernel<kVoiceThreads, P>;                                      \
      FXC_HIP(hipFuncSetAttribute((const void *)k,                                         \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));  \
      hipLaunchKernelGGL(k, grid, dim3(kVoiceThreads), lds, st, Ag, Alin, wts, off, S.dev, \
                         S.tw_f, S.tw_i, nrows, nbatch, nfreq, power, active, gtab, spec); \
    }                                                                                      \
    FXC_HIP(hipGetLastError());                                                            \
    return FXC_OK;                                                                         \
  }
  FXC_VOICE(4)
  FXC_VOICE(8)
  FXC_VOICE(12)
#undef FXC_VOICE
  FXC_CHECK_ARG(false, "fxc_tfpws: length %d too long for the voice kernel", n);
  return FXC_ERR_ARG;
}

static int check_stack_args(const float *pre, const float *wts, const long long *off,
                            const float *lin, int npairs, int n, const float *out) {
  FXC_CHECK_ARG(pre && wts && off && lin && out && npairs >= 1 && n >= 2,
                "stack: invalid arguments");
  FXC_CHECK_ARG(off[0] == 0, "stack: group offsets must start at 0");
  for (int i = 0; i < npairs; ++i)
    FXC_CHECK_ARG(off[i + 1] > off[i], "stack: pair %d has no groups", i);
  return FXC_OK;
}

// Host side of the streaming engine's running PWS (lengths n up to 4096
// positions per 256-thread workgroup, the direct or Bluestein transform).
static int stream_pws_launch_check(int n, StackFft *S, size_t *lds) {
  FXC_TRY(stack_fft_plan(n, S));
  FXC_CHECK_ARG((n + kStkThreads - 1) / kStkThreads <= 16, "stream PWS: length %d too long", n);
  *lds = ((size_t)S->dev.ld + S->twc) * sizeof(float2);
  FXC_CHECK_ARG((size_t)S->dev.ld + S->twc + kLdsReserve <= (size_t)kLdsComplexCapacity,
                "stream PWS: length %d too long for the LDS transform", n);
  return FXC_OK;
}

int stream_pws_close(float2 *S, float *open, const int *d_slot, const int *d_cnt, long long nclose,
                     int n, hipStream_t st, const float *rows) {
  if (nclose == 0) return FXC_OK;
  StackFft F;
  size_t lds = 0;
  FXC_TRY(stream_pws_launch_check(n, &F, &lds));
  const int per = (n + kStkThreads - 1) / kStkThreads;
  // rows per workgroup: as many of kCloseRows as the LDS holds beside the
  // twiddles (3 at n = 2001, where three workgroups share a CU)
  const int nrows_wg = (int)std::min<long long>(
      kCloseRows, ((long long)kLdsComplexCapacity - kLdsReserve - F.twc) / F.dev.ld);
  FXC_CHECK_ARG(nrows_wg >= 1, "stream PWS: length %d does not fit the LDS", n);
  lds = ((size_t)nrows_wg * F.dev.ld + F.twc) * sizeof(float2);
  const unsigned grid = (unsigned)ceil_div((size_t)nclose, (size_t)nrows_wg);
#define FXC_SPC(P)                                                                            \
  if (per <= P) {                                                                            \
    auto *k = F.dev.blue ? stream_pws_close_kernel<P, true> : stream_pws_close_kernel<P, false>; \
    FXC_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                (int)lds));                                                  \
    hipLaunchKernelGGL(k, dim3(grid), dim3(kStkThreads), lds, st, S, open, d_slot, d_cnt, rows, \
                       nclose, nrows_wg, F.dev, F.tw_f, F.tw_i);                             \
    FXC_HIP(hipGetLastError());                                                              \
    return FXC_OK;                                                                           \
  }
  FXC_SPC(4)
  FXC_SPC(8)
  FXC_SPC(16)
#undef FXC_SPC
  return FXC_ERR_ARG;
}

int stream_pws_read(const float2 *S, const float *open, const double *acc, const int *d_slot,
                    const long long *d_total, const int *d_nopen, long long npairs, int n,
                    float power, float *out, hipStream_t st) {
  StackFft F;
  size_t lds = 0;
  FXC_TRY(stream_pws_launch_check(n, &F, &lds));
  const int per = (n + kStkThreads - 1) / kStkThreads;
#define FXC_SPR(P)                                                                            \
  if (per <= P) {                                                                            \
    auto *k = stream_pws_read_kernel<P>;                                                     \
    FXC_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                (int)lds));                                                  \
    hipLaunchKernelGGL(k, dim3((unsigned)npairs), dim3(kStkThreads), lds, st, S, open, acc,  \
                       d_slot, d_total, d_nopen, power, F.dev, F.tw_f, F.tw_i, out);         \
    FXC_HIP(hipGetLastError());                                                              \
    return FXC_OK;                                                                           \
  }
  FXC_SPR(4)
  FXC_SPR(8)
  FXC_SPR(16)
#undef FXC_SPR
  return FXC_ERR_ARG;
}

}  // namespace fxc

using namespace fxc;

// PWS proper over analytic spectra A and device offsets already in place.
static int pws_launch(const StackFft &S, int nbatch, int npairs, const float2 *A,
                      const float *d_weights, const long long *offsets, const float *d_linear,
                      float power, float *d_out, hipStream_t st) {
  const int n = S.dev.n;
  const int per = (n + kVoiceThreads - 1) / kVoiceThreads;
  const size_t lds = ((size_t)nbatch * S.dev.ld + S.twc) * sizeof(float2);
#define FXC_PWS(P)                                                                            \
  if (per <= P) {                                                                            \
    auto *k = pws_kernel<kVoiceThreads, P>;                                                  \
    FXC_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                (int)lds));                                                  \
    hipLaunchKernelGGL(k, dim3((unsigned)npairs), dim3(kVoiceThreads), lds, st, A, d_weights, \
                       offsets, d_linear, S.dev, S.tw_f, S.tw_i, nbatch, power, d_out);      \
    FXC_HIP(hipGetLastError());                                                              \
    return FXC_OK;                                                                           \
  }
  FXC_PWS(4)
  FXC_PWS(8)
  FXC_PWS(16)
#undef FXC_PWS
  FXC_CHECK_ARG(false, "fxc_pws: length %d too long", n);
  return FXC_ERR_ARG;
}

static int check_spec(const fxc_stack_spec *spec, float *power, int *nvoices, int n) {
  *power = 1.f;
  *nvoices = 0;
  if (!spec) return FXC_OK;
  FXC_CHECK_ARG(std::isfinite(spec->power) && spec->power > 0.f,
                "stack: power must be finite and positive");
  FXC_CHECK_ARG(spec->nvoices == 0 || spec->nvoices >= 2, "stack: nvoices must be 0 or >= 2");
  *power = spec->power;
  *nvoices = spec->nvoices >= n / 2 + 1 ? 0 : spec->nvoices;
  return FXC_OK;
}

extern "C" int fxc_pws(const float *d_prestack, const float *d_weights, const long long *h_offsets,
                       const float *d_linear, int npairs, int n, float *d_out, void *stream) {
  return fxc_pws_spec(d_prestack, d_weights, h_offsets, d_linear, npairs, n, nullptr, d_out,
                      stream);
}

extern "C" int fxc_pws_spec(const float *d_prestack, const float *d_weights,
                            const long long *h_offsets, const float *d_linear, int npairs, int n,
                            const fxc_stack_spec *spec, float *d_out, void *stream) {
  FXC_TRY(check_stack_args(d_prestack, d_weights, h_offsets, d_linear, npairs, n, d_out));
  float power = 1.f;
  int nvoices = 0;
  FXC_TRY(check_spec(spec, &power, &nvoices, n));
  StackFft S;
  FXC_TRY(stack_fft_plan(n, &S));
  // the rows must also hold W (n values) after the last batch
  const int nbatch = std::max(stack_batch(S, 0, 0, h_offsets, npairs),
                              (int)ceil_div((size_t)n, (size_t)S.dev.ld));
  FXC_CHECK_ARG((size_t)nbatch * S.dev.ld + S.twc + kLdsReserve <= (size_t)kLdsComplexCapacity,
                "fxc_pws: length %d too long for the LDS transforms", n);
  const int per = (n + kVoiceThreads - 1) / kVoiceThreads;
  FXC_CHECK_ARG(per <= 16, "fxc_pws: length %d too long", n);
  hipStream_t st = (hipStream_t)stream;
  StackWork w;
  FXC_TRY(stack_common(d_prestack, d_linear, h_offsets, npairs, n, w, S, st, false));
  FXC_TRY(pws_launch(S, nbatch, npairs, w.A, d_weights, w.offsets, d_linear, power, d_out, st));
  FXC_HIP(hipStreamSynchronize(st));  // workspace is released on return
  return FXC_OK;
}

// ------------------------------------------------------------------------
// PWS without allocation, for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops).  The workspace of
// fxc_pws (offsets and analytic spectra) lives in a file-scope device array
// of kNoallocScratchBytes per GPU, resolved to a device address once per
// device by fxc_stack_prepare, which also builds the length's FFT plans.
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU race.  Nothing in it has to start at zero (offsets and spectra are
// written in full before they are read), so nothing is cleared.
// ------------------------------------------------------------------------
namespace fxc {
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of `scratch` there

// Bytes of scratch fxc_pws_noalloc needs: offsets, then the spectra on a
// 256-byte boundary.
static size_t noalloc_bytes(long long groups, int npairs, int n, size_t *spec_off) {
  *spec_off = ((size_t)(npairs + 1) * sizeof(long long) + 255) / 256 * 256;
  return *spec_off + (size_t)std::max<long long>(groups, 1) * n * sizeof(float2);
}
}  // namespace fxc

extern "C" int fxc_stack_prepare(int n) {
  FXC_CHECK_ARG(n >= 2, "fxc_stack_prepare: invalid length %d", n);
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  StackFft S;
  FXC_TRY(stack_fft_plan(n, &S, true));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr.count(dev)) {
    void *p = nullptr;
    FXC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FXC_CHECK_ARG(p, "fxc_stack_prepare: no device address for the scratch array");
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  return FXC_OK;
}

extern "C" int fxc_pws_noalloc(const float *d_prestack, const float *d_weights,
                               const long long *h_offsets, const float *d_linear, int npairs,
                               int n, float *d_out, void *stream) {
  FXC_TRY(check_stack_args(d_prestack, d_weights, h_offsets, d_linear, npairs, n, d_out));
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) base = it->second;
  }
  FXC_CHECK_ARG(base, "fxc_pws_noalloc: call fxc_stack_prepare(n) on device %d first", dev);
  StackFft S;
  FXC_TRY(stack_fft_plan(n, &S, false));
  size_t spec_off = 0;
  const size_t need = noalloc_bytes(h_offsets[npairs], npairs, n, &spec_off);
  FXC_CHECK_ARG(need <= kNoallocScratchBytes,
                "fxc_pws_noalloc: %d pairs with %lld groups of length %d need %zu bytes of "
                "workspace, more than the %zu of the fixed scratch array",
                npairs, h_offsets[npairs], n, need, kNoallocScratchBytes);
  const int nbatch = std::max(stack_batch(S, 0, 0, h_offsets, npairs),
                              (int)ceil_div((size_t)n, (size_t)S.dev.ld));
  FXC_CHECK_ARG((size_t)nbatch * S.dev.ld + S.twc + kLdsReserve <= (size_t)kLdsComplexCapacity,
                "fxc_pws_noalloc: length %d too long for the LDS transforms", n);
  FXC_CHECK_ARG((n + kVoiceThreads - 1) / kVoiceThreads <= 16, "fxc_pws_noalloc: length %d too long",
                n);
  hipStream_t st = (hipStream_t)stream;
  long long *offsets = (long long *)base;
  float2 *A = (float2 *)(base + spec_off);
  FXC_TRY(stack_analytic(d_prestack, d_linear, h_offsets, npairs, n, A, offsets, S, st, false));
  FXC_TRY(pws_launch(S, nbatch, npairs, A, d_weights, offsets, d_linear, 1.f, d_out, st));
  // complete on return, as fxc_pws (the host offsets are read by the copy)
  FXC_HIP(hipStreamSynchronize(st));
  return FXC_OK;
}

extern "C" int fxc_tfpws(const float *d_prestack, const float *d_weights,
                         const long long *h_offsets, const float *d_linear, int npairs, int n,
                         int band_limited, float delta, float fmin, float fmax, float taper_hz,
                         float *d_out, void *stream) {
  return fxc_tfpws_spec(d_prestack, d_weights, h_offsets, d_linear, npairs, n, band_limited, delta,
                        fmin, fmax, taper_hz, nullptr, d_out, stream);
}

extern "C" int fxc_tfpws_spec(const float *d_prestack, const float *d_weights,
                              const long long *h_offsets, const float *d_linear, int npairs, int n,
                              int band_limited, float delta, float fmin, float fmax,
                              float taper_hz, const fxc_stack_spec *spec, float *d_out,
                              void *stream) {
  FXC_TRY(check_stack_args(d_prestack, d_weights, h_offsets, d_linear, npairs, n, d_out));
  float power = 1.f;
  int nvoices = 0;
  FXC_TRY(check_spec(spec, &power, &nvoices, n));
  FXC_CHECK_ARG(!band_limited || (delta > 0.f && fmax > fmin), "fxc_tfpws: invalid band");
  StackFft S;
  FXC_TRY(stack_fft_plan(n, &S));
  // rows: up to kStkBatch groups + the linear stack, as many as the LDS holds
  long long gmax = 1;
  for (int i = 0; i < npairs; ++i) gmax = std::max(gmax, h_offsets[i + 1] - h_offsets[i]);
  const long long room = (long long)kLdsComplexCapacity - kLdsReserve - (n + 1) / 2 - S.twc;
  const int nrows = (int)std::min<long long>(std::min<long long>(kStkBatch, gmax) + 1,
                                             room / S.dev.ld);
  FXC_CHECK_ARG(nrows >= 1, "fxc_tfpws: length %d too long for the LDS transforms", n);
  const int nbatch = (int)std::min<long long>(nrows, std::min<long long>(kStkBatch, gmax));
  FXC_CHECK_ARG((n + kVoiceThreads - 1) / kVoiceThreads <= 12,
                "fxc_tfpws: length %d too long for the voice kernel", n);
  hipStream_t st = (hipStream_t)stream;
  StackWork w;
  FXC_TRY(stack_common(d_prestack, d_linear, h_offsets, npairs, n, w, S, st, true));
  const int nfreq = n / 2 + 1;
  const int nchunks = (nfreq + kVoiceChunk - 1) / kVoiceChunk;
  const float df = band_limited ? (float)(1.0 / ((double)n * delta)) : 0.f;
  FXC_TRY(dmalloc((void **)&w.spec, (size_t)npairs * n * sizeof(float2)));
  FXC_HIP(hipMemsetAsync(w.spec, 0, (size_t)npairs * n * sizeof(float2), st));
  if (band_limited) {
    // chunks with no voice inside [fmin, fmax] only carry the linear spectrum
    std::vector<unsigned char> act(nchunks, 0);
    for (int c = 0; c < nchunks; ++c) {
      const int a = c * kVoiceChunk, b = std::min(nfreq, a + kVoiceChunk) - 1;
      const double lo = a * (1.0 / ((double)n * delta)), hi = b * (1.0 / ((double)n * delta));
      act[c] = !(hi < fmin || lo > fmax);
    }
    FXC_TRY(dmalloc((void **)&w.active, nchunks));
    FXC_HIP(hipMemcpyAsync(w.active, act.data(), nchunks, hipMemcpyHostToDevice, st));
  }
  const long long ng = h_offsets[npairs];
  if (nvoices == 0) {
    const int gh = (n + 1) / 2 + 1;
    FXC_TRY(dmalloc((void **)&w.gauss, (size_t)nfreq * gh * sizeof(float)));
    hipLaunchKernelGGL(tfpws_gauss_kernel,
                       dim3((unsigned)ceil_div((size_t)gh, (size_t)kStkThreads), (unsigned)nfreq),
                       dim3(kStkThreads), 0, st, n, nfreq, w.gauss);
    FXC_HIP(hipGetLastError());
  }
  size_t lds = ((size_t)(n + 1) / 2 + (size_t)nrows * S.dev.ld + S.twc) * sizeof(float2);
  FXC_TRY(launch_voice(n, lds, dim3((unsigned)nchunks, (unsigned)npairs), st, w.A,
                       w.A + (size_t)ng * n, d_weights, w.offsets, S, nrows, nbatch, nfreq,
                       power, nvoices, (const unsigned char *)w.active, w.gauss, w.spec));
  lds = ((size_t)S.dev.ld + S.twc) * sizeof(float2);
  FXC_HIP(hipFuncSetAttribute((const void *)tfpws_final_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(tfpws_final_kernel, dim3((unsigned)npairs), dim3(kStkThreads), lds, st, w.spec,
                     w.A + (size_t)ng * n, nfreq, band_limited, df, fmin, fmax, taper_hz, S.dev,
                     S.tw_f, S.tw_i, d_out);
  FXC_HIP(hipGetLastError());
  FXC_HIP(hipStreamSynchronize(st));  // workspace is released on return
  return FXC_OK;
}

extern "C" int fxc_linear_stack(const float *d_records, const long long *h_offsets, int npairs,
                                int n, float *d_out, void *stream) {
  FXC_CHECK_ARG(d_records && h_offsets && d_out && npairs >= 1 && n >= 1,
                "fxc_linear_stack: invalid arguments");
  FXC_CHECK_ARG(h_offsets[0] == 0, "fxc_linear_stack: offsets must start at 0");
  for (int i = 0; i < npairs; ++i)
    FXC_CHECK_ARG(h_offsets[i + 1] > h_offsets[i], "fxc_linear_stack: pair %d has no records", i);
  FXC_CHECK_ARG(npairs <= 65535, "fxc_linear_stack: at most 65535 pairs per call");
  hipStream_t st = (hipStream_t)stream;
  StackWork w;
  FXC_TRY(dmalloc((void **)&w.offsets, (size_t)(npairs + 1) * sizeof(long long)));
  FXC_HIP(hipMemcpyAsync(w.offsets, h_offsets, (size_t)(npairs + 1) * sizeof(long long),
                         hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(linear_stack_kernel,
                     dim3((unsigned)ceil_div((size_t)n, (size_t)kLinThreads), (unsigned)npairs),
                     dim3(kLinThreads), 0, st, d_records, w.offsets, n, d_out);
  FXC_HIP(hipGetLastError());
  FXC_HIP(hipStreamSynchronize(st));  // workspace is released on return
  return FXC_OK;
}
