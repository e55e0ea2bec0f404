// xc_fast hot path: windowed cross-spectrum stack and lag synthesis.
//
// Replaces, for one timestamp and one GPU:
//   * accumulateStepXc2DKernel launched once per window
//     (native/xc/kernels/cuda.kernels.cu:16-36, native/xc/worker/async_workers.cu:66-83):
//       stack[p,k] += conj(S[src_p,k]) * S[rec_p,k] * (1/nstep)
//   * finalize_xc_batch (native/xc/worker/executor.cu:6-45): (-1)^k shift, C2R of
//     length nfft = 2(nspec-1), x 1/(nfft*dt), keep lags [nfft/2-H, nfft/2+H].
//
// MI355X design.  The window sum is a per-bin Hermitian contraction
// C_k = A_k^H B_k over the window axis, so pairs are tiled 32 x 32 stations and
// each workgroup computes one tile for 16 consecutive bins, reading every
// station spectrum once per tile instead of once per pair-window.  Three
// arithmetic forms share the tiling:
//   FP32    plain fp32 FMA on the vector ALU (the parity path);
//   BF16X3  spectra split into bf16 hi+lo, three MFMA products
//           (hi*hi + hi*lo + lo*hi) with fp32 accumulation;
//   FP16    one fp16 MFMA product with a power-of-two scale per station;
//   FP16X3  the scaled spectra split into fp16 hi+lo, three products as BF16X3.
// A fourth form, PLAIN, keeps the reference's own step order on the GPU (the
// stack row of every pair read, updated and written once per window, fp32);
// it exists to time the reference's method beside the tiled forms.
// The cross-spectra go to an HBM workspace in [pair][bin] rows; the lag
// kernel folds each row into an 18522-point complex inverse FFT held wholly
// in LDS and writes only the 2H+1 lags.
#include "fxc_fft_fixed.hpp"
#include "fxc_fft_large.hpp"
#include "fxc_plan_cache.hpp"
#include "fastxc_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <map>
#include <new>
#include <vector>

namespace fxc {

constexpr int kTile = 32;          // stations per tile side
constexpr int kBinChunk = 16;      // bins per contraction workgroup
constexpr int kTilePairs = kTile * kTile;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Spectra split for the MFMA forms.  Input: [nstep][nsta][nspec] complex64
// (the StepPack payload order).  Output per (station, bin): K' = 2*Kh 16-bit
// values [Re w=0..Kh-1 | Im w=0..Kh-1] (Kh = nstep rounded up to 16),
// windows >= nstep zero, in
// [nsta][nspec][K'] order, for hi and (BF16X3) lo parts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short bf16_bits_rne(float x) {
  unsigned int u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf/nan kept
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_bits_to_float(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}

__global__ void station_absmax_kernel(const float2 *__restrict__ spec, int nstep, int nsta,
                                      int nspec, unsigned int *__restrict__ amax_bits) {
  const int s = blockIdx.y;
  float m = 0.f;
  for (int w = 0; w < nstep; ++w) {
    const float2 *row = spec + ((size_t)w * nsta + s) * nspec;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nspec; k += gridDim.x * blockDim.x) {
      const float2 v = row[k];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
  }
  // wave reduction then one atomic per wave (positive floats order as ints)
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(amax_bits + s, __float_as_uint(m));
}

// scale_s = 2^-e with 2^(e-1) <= max < 2^e, so |S/scale| <= 1 (exact power).
__global__ void station_scale_kernel(const unsigned int *__restrict__ amax_bits, int nsta,
                                     float *__restrict__ inv_scale, float *__restrict__ scale) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsta) return;
  const float m = __uint_as_float(amax_bits[s]);
  int e = 0;
  if (m > 0.f && isfinite(m)) (void)frexpf(m, &e);
  inv_scale[s] = ldexpf(1.f, -e);
  scale[s] = ldexpf(1.f, e);
}

// Split pass of the MFMA forms.  Operand layout [bin][K group][station][8]:
// K' = 2*kh values per (station, bin) -- window w's real part at K index 2w,
// its imaginary part at 2w + 1, zero for w >= nstep -- cut into groups of 8
// consecutive K indices (16 bytes, 4 windows), with the stations of one
// group adjacent,
// so the 16 lanes that hold one K group of 16 stations in an MFMA operand
// read 256 contiguous bytes.  One workgroup converts 64 stations x 16 bins:
// each block of 8 windows is staged through LDS so the reads run along the
// bins and the writes along the stations.
constexpr int kSplitSta = 64;
constexpr int kSplitBins = 16;
constexpr int kSplitThreads = 256;
constexpr int kSplitLoads = 8 * kSplitSta * kSplitBins / kSplitThreads;  // per thread and block

// One workgroup's part: source stations [s0, s0 + 64) of spectra
// [nstep][nsta][nspec], bins [k0, k0 + 16).  Source station s lands at operand
// row sta_off + s of an operand with nsta_dst rows; station-windows whose
// valid byte ([nstep][nsta], when given) is 0 enter as zeros.
template <int kMode>
__device__ __forceinline__ void split_part(const float2 *__restrict__ spec, int nstep, int nsta,
                                           int nspec, int kh, const float *__restrict__ inv_scale,
                                           const unsigned char *__restrict__ valid, int sta_off,
                                           int nsta_dst, int s0, int k0,
                                           unsigned short *__restrict__ hi,
                                           unsigned short *__restrict__ lo) {
  __shared__ float2 tile[8][kSplitSta][kSplitBins + 1];
  const int tid = threadIdx.x;
  const int ngrp = kh / 4;  // K groups of 8 values: kh/8 real + kh/8 imaginary
  for (int wb = 0; wb < kh; wb += 8) {
    // load windows wb..wb+7: 16 lanes read one station's 16 bins (128 bytes);
    // every load of the block is issued before the first LDS store, so their
    // latencies overlap
    float2 z[kSplitLoads];
#pragma unroll
    for (int j = 0; j < kSplitLoads; ++j) {
      const int i = tid + j * kSplitThreads;
      const int kk = i % kSplitBins;
      const int ss = (i / kSplitBins) % kSplitSta;
      const int ww = i / (kSplitBins * kSplitSta);
      const int w = wb + ww, sidx = s0 + ss, k = k0 + kk;
      z[j] = make_float2(0.f, 0.f);
      if (w < nstep && sidx < nsta && k < nspec && (!valid || valid[(size_t)w * nsta + sidx]))
        z[j] = spec[((size_t)w * nsta + sidx) * nspec + k];
    }
    __syncthreads();  // the previous block's conversion is done with the tile
#pragma unroll
    for (int j = 0; j < kSplitLoads; ++j) {
      const int i = tid + j * kSplitThreads;
      tile[i / (kSplitBins * kSplitSta)][(i / kSplitBins) % kSplitSta][i % kSplitBins] = z[j];
    }
    __syncthreads();
    // convert: thread = (station, bin); 64 consecutive threads write 64
    // consecutive stations' 16-byte groups
    for (int i = tid; i < kSplitSta * kSplitBins; i += kSplitThreads) {
      const int ss = i % kSplitSta;
      const int kk = i / kSplitSta;
      const int sidx = s0 + ss, k = k0 + kk;
      if (sidx >= nsta || k >= nspec) continue;
      const float sc = (kMode == FXC_PREC_FP16 || kMode == FXC_PREC_FP16X3)
                           ? inv_scale[sta_off + sidx]
                           : 1.f;
#pragma unroll
      for (int part = 0; part < 2; ++part) {  // windows wb + 4 part .. + 3
        unsigned short h[8], l[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const float2 z = tile[4 * part + j / 2][ss][kk];
          const float v = ((j & 1) ? z.y : z.x) * sc;
          if (kMode == FXC_PREC_FP16) {
            h[j] = __builtin_bit_cast(unsigned short, (_Float16)v);
            l[j] = 0;
          } else if (kMode == FXC_PREC_FP16X3) {
            const _Float16 hh = (_Float16)v;
            h[j] = __builtin_bit_cast(unsigned short, hh);
            l[j] = __builtin_bit_cast(unsigned short, (_Float16)(v - (float)hh));
          } else {
            const unsigned short hb = bf16_bits_rne(v);
            h[j] = hb;
            l[j] = bf16_bits_rne(v - bf16_bits_to_float(hb));
          }
        }
        const int g = 2 * (wb / 8) + part;
        const size_t o = (((size_t)k * ngrp + g) * nsta_dst + sta_off + sidx) * 8;
        uint4 hv;
        hv.x = h[0] | ((unsigned)h[1] << 16); hv.y = h[2] | ((unsigned)h[3] << 16);
        hv.z = h[4] | ((unsigned)h[5] << 16); hv.w = h[6] | ((unsigned)h[7] << 16);
        *(uint4 *)(hi + o) = hv;
        if (kMode == FXC_PREC_BF16X3 || kMode == FXC_PREC_FP16X3) {
          uint4 lv;
          lv.x = l[0] | ((unsigned)l[1] << 16); lv.y = l[2] | ((unsigned)l[3] << 16);
          lv.z = l[4] | ((unsigned)l[5] << 16); lv.w = l[6] | ((unsigned)l[7] << 16);
          *(uint4 *)(lo + o) = lv;
        }
      }
    }
  }
}

template <int kMode>
__global__ void __launch_bounds__(kSplitThreads)
split_kernel(const float2 *__restrict__ spec, int nstep, int nsta, int nspec, int kh,
             const float *__restrict__ inv_scale, unsigned short *__restrict__ hi,
             unsigned short *__restrict__ lo, int bin0) {
  split_part<kMode>(spec, nstep, nsta, nspec, kh, inv_scale, nullptr, 0, nsta,
                    blockIdx.y * kSplitSta, bin0 + blockIdx.x * kSplitBins, hi, lo);
}

// Uneven segments in one launch: segment z = blockIdx.z holds spectra
// [nstep][nsta][nspec] at spec_off complex values and valid bytes [nstep][nsta]
// at valid_off; its stations take operand rows [sta_off, sta_off + nsta).
struct SplitSeg {
  long long spec_off;
  long long valid_off;
  int nsta;
  int nstep;
  int sta_off;
  int npairs;
};

template <int kMode>
__global__ void __launch_bounds__(kSplitThreads)
split_segments_kernel(const float2 *__restrict__ spec, const unsigned char *__restrict__ valid,
                      const SplitSeg *__restrict__ segs, int nsta_dst, int nspec, int kh,
                      unsigned short *__restrict__ hi, unsigned short *__restrict__ lo) {
  const SplitSeg sg = segs[blockIdx.z];
  const int s0 = blockIdx.y * kSplitSta;
  if (s0 >= sg.nsta) return;  // the whole workgroup leaves before any barrier
  split_part<kMode>(spec + sg.spec_off, sg.nstep, sg.nsta, nspec, kh, nullptr,
                    valid ? valid + sg.valid_off : nullptr, sg.sta_off, nsta_dst, s0,
                    blockIdx.x * kSplitBins, hi, lo);
}

// ---------------------------------------------------------------------------
// Tile descriptors: tile t covers src stations [src_base, +32) and rec
// stations [rec_base, +32); rows[t][i][j] is the workspace row of pair
// (src_base+i, rec_base+j) or -1.
// ---------------------------------------------------------------------------
struct TileDesc {
  int src_base;
  int rec_base;
};

// Streaming store of one complex value: the cross-spectra are written once
// and read back by another kernel, so they must not evict the station
// spectra that the contraction re-reads from L2.
__device__ __forceinline__ void store_stream(float2 *dst, float2 v) {
  __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v), (unsigned long long *)dst);
}
// Two adjacent values (16 bytes; dst 16-byte aligned).
__device__ __forceinline__ void store_stream2(float2 *dst, float2 a, float2 b) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  u64x2 v;
  v.x = __builtin_bit_cast(unsigned long long, a);
  v.y = __builtin_bit_cast(unsigned long long, b);
  __builtin_nontemporal_store(v, (u64x2 *)dst);
}

// Work item -> (tile, bin chunk) of the FP32 contraction.  Workgroups are
// dispatched round-robin over the 8 XCDs, so 8 consecutive items take 8
// consecutive bin chunks of one tile: each XCD's L2 then serves one chunk of
// every station's spectra, and each pair row receives 8 x 128 contiguous
// bytes at once.  (The MFMA contraction keeps the same chunk-per-XCD rule
// with its own order of tiles; see contract_mfma_kernel.)
constexpr int kChunkGroup = 8;

__device__ __forceinline__ bool tile_chunk_of(long long item, int ntiles, int nchunks, int &tile,
                                              int &chunk) {
  const long long per_group = (long long)ntiles * kChunkGroup;
  const int g = (int)(item / per_group);
  const int r = (int)(item - g * per_group);
  tile = r / kChunkGroup;
  chunk = g * kChunkGroup + (r - tile * kChunkGroup);
  return chunk < nchunks;
}
__device__ __forceinline__ bool tile_chunk(int ntiles, int nchunks, int &tile, int &chunk) {
  return tile_chunk_of(blockIdx.x, ntiles, nchunks, tile, chunk);
}

// FP32 contraction: thread = (bin, 8x8 micro-tile of pairs).
__global__ void __launch_bounds__(256)
contract_fp32_kernel(const float2 *__restrict__ spec, int nstep, int nsta, int nspec,
                     const TileDesc *__restrict__ tiles, const int *__restrict__ tile_rows,
                     float2 *__restrict__ xspec, long long xpitch, int ntiles, float scale,
                     int chunk0, int nchunks) {
  __shared__ int rows[kTilePairs];
  __shared__ int active[16];
  int tile, chunk;
  if (!tile_chunk(ntiles, nchunks, tile, chunk)) return;
  chunk += chunk0;
  const int tid = threadIdx.x;
  const int *trow = tile_rows + (size_t)tile * kTilePairs;
  for (int i = tid; i < kTilePairs; i += blockDim.x) rows[i] = trow[i];
  if (tid < 16) active[tid] = 0;
  __syncthreads();
  for (int i = tid; i < kTilePairs; i += blockDim.x) {
    if (rows[i] >= 0) {
      const int mi = (i >> 5) >> 3, mj = (i & 31) >> 3;
      active[mi * 4 + mj] = 1;
    }
  }
  __syncthreads();
  const int b = tid & 15;
  const int g = tid >> 4;
  const int mi = g >> 2, mj = g & 3;
  const int k = chunk * kBinChunk + b;
  if (k >= nspec || !active[g]) return;
  const TileDesc td = tiles[tile];
  const int s0 = td.src_base + mi * 8;
  const int r0 = td.rec_base + mj * 8;
  float2 acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[a][c] = make_float2(0.f, 0.f);
  // software pipeline: loads of window w+1 are in flight while window w is
  // accumulated
  const size_t wstride = (size_t)nsta * nspec;
  float2 sv[8], rv[8], sn[8], rn[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    sv[a] = (s0 + a < nsta) ? spec[(size_t)(s0 + a) * nspec + k] : make_float2(0.f, 0.f);
    rv[a] = (r0 + a < nsta) ? spec[(size_t)(r0 + a) * nspec + k] : make_float2(0.f, 0.f);
  }
  for (int w = 0; w < nstep; ++w) {
    if (w + 1 < nstep) {
      const float2 *base = spec + (size_t)(w + 1) * wstride + k;
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        sn[a] = (s0 + a < nsta) ? base[(size_t)(s0 + a) * nspec] : make_float2(0.f, 0.f);
        rn[a] = (r0 + a < nsta) ? base[(size_t)(r0 + a) * nspec] : make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        // conj(s) * r
        acc[a][c].x += sv[a].x * rv[c].x + sv[a].y * rv[c].y;
        acc[a][c].y += sv[a].x * rv[c].y - sv[a].y * rv[c].x;
      }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      sv[a] = sn[a];
      rv[a] = rn[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 8; ++a) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int row = rows[(mi * 8 + a) * kTile + mj * 8 + c];
      if (row >= 0) store_stream(xspec + (size_t)row * xpitch + k, cscale(acc[a][c], scale));
    }
  }
}

// MFMA contraction.  Per bin: Cr = A.B1, Ci = A.B2 with a_s = [Re | Im] and
// b2_s = [Im | -Re], so conj(a)*b = Cr + i Ci.  Results are staged in LDS as
// [pair][bin] and written as 128-byte rows.
// Operand fragments of one lane for one bin and one split part (hi or lo):
// K' = 2*kh = 16*NK16 with kh a multiple of 8: NK16/2 K steps of one
// 16x16x32 MFMA each, and for odd NK16 a last step of one 16x16x16 MFMA, so
// a window count such as 24 is not padded to 32.
// Source side: A = (Re, Im) per window; receiver side B1 = (Re, Im) and
// B2 = (Im, -Re) per window, so A.B1 = Re(conj(a) b), A.B2 = Im(conj(a) b).
// B2 is formed in registers from B1: each 32-bit word holds one window's
// (Re, Im), and B2's word is that word rotated by 16 bits with the new high
// half negated.
template <int NK16>
struct SrcFrag {
  static constexpr int kFull = NK16 / 2;
  static constexpr bool kHalf = (NK16 & 1) != 0;
  uint4 a32[kFull > 0 ? kFull : 1];
  uint2 a16;  // the 16-deep last step (kHalf)
};
template <int NK16>
struct RecFrag {
  static constexpr int kFull = NK16 / 2;
  static constexpr bool kHalf = (NK16 & 1) != 0;
  uint4 b32[kFull > 0 ? kFull : 1];
  uint2 b16;
};
// B2 of one receiver fragment.
template <int NK16>
struct RecB2 {
  uint4 c32[RecFrag<NK16>::kFull > 0 ? RecFrag<NK16>::kFull : 1];
  uint2 c16;
};

// (Re, Im) -> (Im, -Re) in every 32-bit word of 16-bit values.
__device__ __forceinline__ unsigned swap_neg(unsigned d) {
  return ((d >> 16) | (d << 16)) ^ 0x80000000u;
}
__device__ __forceinline__ uint4 swap_neg(uint4 v) {
  return make_uint4(swap_neg(v.x), swap_neg(v.y), swap_neg(v.z), swap_neg(v.w));
}
__device__ __forceinline__ uint2 swap_neg(uint2 v) { return make_uint2(swap_neg(v.x), swap_neg(v.y)); }

template <int NK16>
__device__ __forceinline__ void derive_b2(const RecFrag<NK16> &y, RecB2<NK16> &z) {
#pragma unroll
  for (int s = 0; s < RecFrag<NK16>::kFull; ++s) z.c32[s] = swap_neg(y.b32[s]);
  if constexpr (RecFrag<NK16>::kHalf) z.c16 = swap_neg(y.b16);
}

// a / b point at K group 0 of one (bin, station) in the split layout; K
// group g (8 values) sits gstride = 8*nsta elements further.  Full K step s
// of lane group lane_g is K group 4s + lane_g; in the 16-deep last step lane
// group lane_g holds 4 values, the half (lane_g & 1) of K group
// 4*kFull + lane_g/2.
template <int NK16>
__device__ __forceinline__ void load_src(const unsigned short *a, size_t gstride, int lane_g,
                                         SrcFrag<NK16> &o) {
  constexpr int kFull = SrcFrag<NK16>::kFull;
#pragma unroll
  for (int s = 0; s < kFull; ++s) o.a32[s] = *(const uint4 *)(a + (4 * s + lane_g) * gstride);
  if constexpr (SrcFrag<NK16>::kHalf)
    o.a16 = *(const uint2 *)(a + (4 * kFull + (lane_g >> 1)) * gstride + (lane_g & 1) * 4);
}

template <int NK16>
__device__ __forceinline__ void load_rec(const unsigned short *b, size_t gstride, int lane_g,
                                         RecFrag<NK16> &o) {
  constexpr int kFull = RecFrag<NK16>::kFull;
#pragma unroll
  for (int s = 0; s < kFull; ++s) o.b32[s] = *(const uint4 *)(b + (4 * s + lane_g) * gstride);
  if constexpr (RecFrag<NK16>::kHalf)
    o.b16 = *(const uint2 *)(b + (4 * kFull + (lane_g >> 1)) * gstride + (lane_g & 1) * 4);
}

template <int kMode>
__device__ __forceinline__ void mfma32(uint4 a, uint4 b, f32x4_t &acc) {
  if constexpr (kMode == FXC_PREC_FP16 || kMode == FXC_PREC_FP16X3) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a),
                                                 __builtin_bit_cast(f16x8_t, b), acc, 0, 0, 0);
  } else {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                  __builtin_bit_cast(bf16x8_t, b), acc, 0, 0, 0);
  }
}

template <int kMode>
__device__ __forceinline__ void mfma16(uint2 a, uint2 b, f32x4_t &acc) {
  if constexpr (kMode == FXC_PREC_FP16 || kMode == FXC_PREC_FP16X3) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(f16x4_t, a),
                                                __builtin_bit_cast(f16x4_t, b), acc, 0, 0, 0);
  } else {
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4_t, a),
                                                    __builtin_bit_cast(short4_t, b), acc, 0, 0, 0);
  }
}

// cr += A.B1, ci += A.B2 over the full K steps of one bin; the 16-deep last
// step accumulates into hr, hi.  The two shapes never share an accumulator
// chain: a 16x16x16 MFMA that took its accumulator straight from a 16x16x32
// MFMA (issued back to back) gave wrong sums, so the 16-deep sums are kept
// apart and added on the VALU.
template <int kMode, int NK16>
__device__ __forceinline__ void mfma_bin(const SrcFrag<NK16> &x, const RecFrag<NK16> &y,
                                         const RecB2<NK16> &y2, f32x4_t &cr, f32x4_t &ci,
                                         f32x4_t &hr, f32x4_t &hi) {
#pragma unroll
  for (int s = 0; s < SrcFrag<NK16>::kFull; ++s) {
    mfma32<kMode>(x.a32[s], y.b32[s], cr);
    mfma32<kMode>(x.a32[s], y2.c32[s], ci);
  }
  if constexpr (SrcFrag<NK16>::kHalf) {
    mfma16<kMode>(x.a16, y.b16, hr);
    mfma16<kMode>(x.a16, y2.c16, hi);
  }
}

// LDS slot of (pair p, bin b) in the [pair][16 bins] staging tile.  The bin
// is XOR-swizzled with the pair's low four bits: an MFMA result register
// holds one bin for 16 consecutive pairs, which would otherwise sit 128 bytes
// apart in one bank group; swizzled, the 16 writes land in 16 distinct bank
// pairs, and a 16-thread group still reads one pair's 16 bins as a
// permutation of one 128-byte row.
__device__ __forceinline__ int stage_index(int p, int b) { return p * kBinChunk + (b ^ (p & 15)); }

// Operands of one bin for the whole 32x32 tile: two 16-station source blocks
// and two 16-station receiver blocks, hi (and lo for the split form).
template <int NK16, bool kSplit>
struct TileOps {
  SrcFrag<NK16> sh[2], sl[kSplit ? 2 : 1];
  RecFrag<NK16> rh[2], rl[kSplit ? 2 : 1];
};
// The source half and the receiver half of TileOps, held apart where
// consecutive items share their sources.
template <int NK16, bool kSplit>
struct SrcOps {
  SrcFrag<NK16> sh[2], sl[kSplit ? 2 : 1];
};
template <int NK16, bool kSplit>
struct RecOps {
  RecFrag<NK16> rh[2], rl[kSplit ? 2 : 1];
};

// Per-lane addressing of one work item (tile x 16-bin chunk).
struct ItemAddr {
  size_t s_off[2], r_off[2];  // station offsets (elements) of the lane's rows/columns
  float col_scale[2];
  float row_scale[2][4];
  int k0;
};

template <int kMode>
__device__ __forceinline__ void item_addr(const TileDesc &td, int chunk, int nsta, int lane_r,
                                          int lane_g, float scale, const float *sta_scale,
                                          ItemAddr &ia) {
  // Stations past the array end load a valid station; their rows/columns
  // belong to pairs that are never stored.
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int s_sta = td.src_base + h * 16 + lane_r;
    const int r_sta = td.rec_base + h * 16 + lane_r;
    ia.s_off[h] = (size_t)(s_sta < nsta ? s_sta : nsta - 1) * 8;
    ia.r_off[h] = (size_t)(r_sta < nsta ? r_sta : nsta - 1) * 8;
    ia.col_scale[h] = scale;
    if constexpr (kMode == FXC_PREC_FP16 || kMode == FXC_PREC_FP16X3)
      ia.col_scale[h] *= (r_sta < nsta ? sta_scale[r_sta] : 0.f);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ia.row_scale[h][r] = 1.f;
      if constexpr (kMode == FXC_PREC_FP16 || kMode == FXC_PREC_FP16X3) {
        const int sa = td.src_base + h * 16 + 4 * lane_g + r;
        ia.row_scale[h][r] = sa < nsta ? sta_scale[sa] : 0.f;
      }
    }
  }
  ia.k0 = chunk * kBinChunk;
}

// kp_total: the operand depth K' of one bin in the split layout (16 NK16,
// or a multiple of it when the windows are taken in chunks of K' = 16 NK16);
// kc: the chunk loaded.
template <int kMode, int NK16, bool kSplit>
__device__ __forceinline__ void load_tile_ops(const unsigned short *__restrict__ hi,
                                              const unsigned short *__restrict__ lo,
                                              const ItemAddr &ia, int b, int nsta, int nspec,
                                              int lane_g, TileOps<NK16, kSplit> &o,
                                              int kp_total = 16 * NK16, int kc = 0) {
  const size_t gstride = (size_t)nsta * 8;
  int k = ia.k0 + b;
  k = k < nspec ? k : nspec - 1;  // tail bins compute a valid bin, never stored
  // bin k: K' x nsta elements; chunk kc starts 2 NK16 K groups further per chunk
  const size_t kb = (size_t)k * kp_total * nsta + (size_t)kc * (2 * NK16) * gstride;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    load_src(hi + kb + ia.s_off[h], gstride, lane_g, o.sh[h]);
    load_rec(hi + kb + ia.r_off[h], gstride, lane_g, o.rh[h]);
    if constexpr (kSplit) {
      load_src(lo + kb + ia.s_off[h], gstride, lane_g, o.sl[h]);
      load_rec(lo + kb + ia.r_off[h], gstride, lane_g, o.rl[h]);
    }
  }
}

// Operands of bin b of an item: the source or the receiver fragments only.
template <int kMode, int NK16, bool kSplit>
__device__ __forceinline__ void load_src_ops(const unsigned short *__restrict__ hi,
                                             const unsigned short *__restrict__ lo,
                                             const ItemAddr &ia, int b, int nsta, int nspec,
                                             int lane_g, SrcOps<NK16, kSplit> &o) {
  const size_t gstride = (size_t)nsta * 8;
  int k = ia.k0 + b;
  k = k < nspec ? k : nspec - 1;  // tail bins compute a valid bin, never stored
  const size_t kb = (size_t)k * (16 * NK16) * nsta;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    load_src(hi + kb + ia.s_off[h], gstride, lane_g, o.sh[h]);
    if constexpr (kSplit) load_src(lo + kb + ia.s_off[h], gstride, lane_g, o.sl[h]);
  }
}
template <int kMode, int NK16, bool kSplit>
__device__ __forceinline__ void load_rec_ops(const unsigned short *__restrict__ hi,
                                             const unsigned short *__restrict__ lo,
                                             const ItemAddr &ia, int b, int nsta, int nspec,
                                             int lane_g, RecOps<NK16, kSplit> &o) {
  const size_t gstride = (size_t)nsta * 8;
  int k = ia.k0 + b;
  k = k < nspec ? k : nspec - 1;
  const size_t kb = (size_t)k * (16 * NK16) * nsta;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    load_rec(hi + kb + ia.r_off[h], gstride, lane_g, o.rh[h]);
    if constexpr (kSplit) load_rec(lo + kb + ia.r_off[h], gstride, lane_g, o.rl[h]);
  }
}

// One bin of the tile from its source fragments (sh, sl) and receiver
// fragments (rh, rl), staged in LDS.
template <int kMode, int NK16, bool kSplit>
__device__ __forceinline__ void compute_tile_bin(const SrcFrag<NK16> (&sh)[2],
                                                 const SrcFrag<NK16> (&sl)[kSplit ? 2 : 1],
                                                 const RecFrag<NK16> (&rh)[2],
                                                 const RecFrag<NK16> (&rl)[kSplit ? 2 : 1],
                                                 const ItemAddr &ia, int b, int lane_r, int lane_g,
                                                 float2 *stage) {
  // every B2 of the bin is formed before the first MFMA: a VALU result that
  // an MFMA reads right away costs wait states
  RecB2<NK16> b2h[2], b2l[kSplit ? 2 : 1];
#pragma unroll
  for (int sj = 0; sj < 2; ++sj) {
    derive_b2(rh[sj], b2h[sj]);
    if constexpr (kSplit) derive_b2(rl[sj], b2l[sj]);
  }
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
      f32x4_t cr = {0.f, 0.f, 0.f, 0.f};
      f32x4_t ci = {0.f, 0.f, 0.f, 0.f};
      f32x4_t hr = {0.f, 0.f, 0.f, 0.f};
      f32x4_t hi = {0.f, 0.f, 0.f, 0.f};
      mfma_bin<kMode>(sh[si], rh[sj], b2h[sj], cr, ci, hr, hi);  // hi.hi
      if constexpr (kSplit) {
        mfma_bin<kMode>(sl[si], rh[sj], b2h[sj], cr, ci, hr, hi);  // lo.hi
        mfma_bin<kMode>(sh[si], rl[sj], b2l[sj], cr, ci, hr, hi);  // hi.lo
      }
      if constexpr (SrcFrag<NK16>::kHalf) {
        if constexpr (SrcFrag<NK16>::kFull > 0) {
          cr += hr;
          ci += hi;
        } else {
          cr = hr;
          ci = hi;
        }
      }
      // D layout: lane holds rows 4*lane_g + r, column lane_r of the 16x16 block.
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = si * 16 + 4 * lane_g + r;
        const int c = sj * 16 + lane_r;
        const float sc = ia.col_scale[sj] * ia.row_scale[si][r];
        stage[stage_index(a * kTile + c, b)] = make_float2(cr[r] * sc, ci[r] * sc);
      }
    }
}

template <int kMode, int NK16, bool kSplit>
__device__ __forceinline__ void compute_tile_bin(const TileOps<NK16, kSplit> &o, const ItemAddr &ia,
                                                 int b, int lane_r, int lane_g, float2 *stage) {
  compute_tile_bin<kMode, NK16, kSplit>(o.sh, o.sl, o.rh, o.rl, ia, b, lane_r, lane_g, stage);
}

// More than 128 windows: the operand depth runs in kchunks chunks of
// K' = 16 NK16 (32 windows at NK16 4), the four blocks' sums held in
// registers across the chunks, then staged as compute_tile_bin stages them.
template <int kMode, int NK16, bool kSplit>
__device__ __forceinline__ void compute_tile_bin_chunked(const unsigned short *__restrict__ hi,
                                                         const unsigned short *__restrict__ lo,
                                                         const ItemAddr &ia, int b, int nsta,
                                                         int nspec, int lane_r, int lane_g,
                                                         int kchunks, float2 *stage) {
  static_assert((NK16 & 1) == 0, "chunks are whole 16x16x32 steps");
  f32x4_t cr[2][2], ci[2][2];
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
      cr[si][sj] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      ci[si][sj] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    }
  const int kp_total = 16 * NK16 * kchunks;
#pragma unroll 1
  for (int kc = 0; kc < kchunks; ++kc) {
    TileOps<NK16, kSplit> o;
    load_tile_ops<kMode>(hi, lo, ia, b, nsta, nspec, lane_g, o, kp_total, kc);
    __builtin_amdgcn_sched_barrier(0);
    RecB2<NK16> b2h[2], b2l[kSplit ? 2 : 1];
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
      derive_b2(o.rh[sj], b2h[sj]);
      if constexpr (kSplit) derive_b2(o.rl[sj], b2l[sj]);
    }
#pragma unroll
    for (int si = 0; si < 2; ++si)
#pragma unroll
      for (int sj = 0; sj < 2; ++sj) {
        f32x4_t hr = {0.f, 0.f, 0.f, 0.f}, hi16 = {0.f, 0.f, 0.f, 0.f};  // unused (no 16-deep step)
        mfma_bin<kMode>(o.sh[si], o.rh[sj], b2h[sj], cr[si][sj], ci[si][sj], hr, hi16);
        if constexpr (kSplit) {
          mfma_bin<kMode>(o.sl[si], o.rh[sj], b2h[sj], cr[si][sj], ci[si][sj], hr, hi16);
          mfma_bin<kMode>(o.sh[si], o.rl[sj], b2l[sj], cr[si][sj], ci[si][sj], hr, hi16);
        }
      }
  }
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = si * 16 + 4 * lane_g + r;
        const int c = sj * 16 + lane_r;
        const float sc = ia.col_scale[sj] * ia.row_scale[si][r];
        stage[stage_index(a * kTile + c, b)] =
            make_float2(cr[si][sj][r] * sc, ci[si][sj][r] * sc);
      }
}

// One work item = one 32x32 station tile x 16 bins; 8 waves, each computing
// the whole tile (four 16x16 MFMA blocks) for two of the 16 bins, so every
// station fragment is read once per workgroup.  Persistent: a workgroup walks
// runs of neighbouring tiles at one bin chunk (the order is set out at the
// top of the kernel), and each wave keeps the source operands of its two
// bins in registers while consecutive tiles share their source stations.
// With small K each wave software-pipelines its bins: the receiver operands
// of the next bin -- of this item or of the next one -- are loaded before the
// current bin's MFMAs, so load latency overlaps the MFMAs, the LDS staging
// and the previous item's stores.
// The LDS staging tile admits one workgroup (8 waves) per CU, 2 waves per
// SIMD; waves_per_eu tells the scheduler so, and it then keeps a bin's
// operand loads in flight together instead of trimming registers for an
// occupancy the LDS never allows.
template <int kMode, int NK16, bool kChunked = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
contract_mfma_kernel(const unsigned short *__restrict__ hi, const unsigned short *__restrict__ lo,
                     int nsta, int nspec, const float *__restrict__ sta_scale,
                     const TileDesc *__restrict__ tiles, const int *__restrict__ tile_rows,
                     float2 *__restrict__ xspec, long long xpitch, int ntiles, float scale,
                     int kchunks, int chunk0, int nchunks) {
  extern __shared__ __align__(16) float2 stage[];  // [kTilePairs][kBinChunk], swizzled
  __shared__ int rows[kTilePairs];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int lane_r = lane & 15, lane_g = lane >> 4;
  constexpr bool kSplit = kMode == FXC_PREC_BF16X3 || kMode == FXC_PREC_FP16X3;
  constexpr bool kPipe = NK16 * (kSplit ? 2 : 1) <= 8;
  constexpr int kBins = kBinChunk / 8;  // bins per wave and item
  static_assert(kBins == 2, "two bins per wave and item");
  // Item order.  Workgroups are dispatched round-robin over the 8 XCDs, so
  // workgroup w runs on XCD x = w mod 8 and takes chunks 8 g + x only: each
  // XCD's L2 then serves one chunk of every station's spectra at a time.
  // Within an XCD's stream j = 0, 1, ... (tile j mod ntiles of chunk group
  // j / ntiles) the U = gridDim.x / 8 workgroups of the XCD take runs of
  // kRun consecutive j, so a workgroup's consecutive items are mostly
  // neighbouring tiles of one tile row at one chunk, which share their
  // source stations.  (gridDim.x is a multiple of 8.)
  constexpr int kRun = 8;
  const int xcd = blockIdx.x & 7, wu = blockIdx.x >> 3, nu = gridDim.x >> 3;
  const long long nj = (long long)ntiles * ((nchunks + kChunkGroup - 1) / kChunkGroup);
  // next valid item at or after this workgroup's k-th (items whose chunk is
  // past the end are skipped); -1 when none is left
  auto next_item = [&](long long k, int &tile, int &chunk) -> long long {
    for (;; ++k) {
      const long long j = (k / kRun) * nu * kRun + (long long)wu * kRun + k % kRun;
      if (j >= nj) return -1;
      const long long g = j / ntiles;
      tile = (int)(j - g * ntiles);
      chunk = (int)(g * kChunkGroup) + xcd;
      if (chunk < nchunks) return k;
    }
  };
  int tile = 0, chunk = 0;
  long long item = next_item(0, tile, chunk);
  if (item < 0) return;
  ItemAddr ia;
  item_addr<kMode>(tiles[tile], chunk0 + chunk, nsta, lane_r, lane_g, scale, sta_scale, ia);
  // (the chunked-depth form takes the unpipelined path)
  constexpr bool pipe = kPipe && !kChunked;
  // unpipelined depths up to 48 windows (40 in split FP16, whose scales
  // take the registers the 48-window sources would need) hold the sources
  // of both bins too
  constexpr bool kHoldSrc = !pipe && !kChunked && NK16 <= (kMode == FXC_PREC_FP16X3 ? 5 : 6);
  // pipelined operands: the sources of the wave's two bins (sa: bin wave,
  // sb: bin wave + 8) stay in registers while consecutive items share their
  // source stations and chunk; the receivers of the next bin load while the
  // current bin's MFMAs run
  SrcOps<NK16, kSplit> sa, sb;
  RecOps<NK16, kSplit> rc;
  bool sb_stale = true;  // sb (and sa, unpipelined) do not hold this item's sources yet
  if constexpr (pipe) {
    load_src_ops<kMode>(hi, lo, ia, wave, nsta, nspec, lane_g, sa);
    load_rec_ops<kMode>(hi, lo, ia, wave, nsta, nspec, lane_g, rc);
  }
  // workspace rows of the tile's 1024 pairs, two per thread, loaded one item
  // ahead so the LDS fill after the barrier does not wait on memory
  static_assert(kTilePairs == 2 * 512, "two row entries per thread");
  int rreg0 = tile_rows[(size_t)tile * kTilePairs + tid];
  int rreg1 = tile_rows[(size_t)tile * kTilePairs + tid + 512];
  while (item >= 0) {
    __syncthreads();  // the previous item's store phase is done with rows/stage
    rows[tid] = rreg0;
    rows[tid + 512] = rreg1;
    int ntile = 0, nchunk = 0;
    const long long nitem = next_item(item + 1, ntile, nchunk);
    ItemAddr na = ia;
    bool same_src = false;
    if (nitem >= 0) {
      item_addr<kMode>(tiles[ntile], chunk0 + nchunk, nsta, lane_r, lane_g, scale, sta_scale, na);
      rreg0 = tile_rows[(size_t)ntile * kTilePairs + tid];
      rreg1 = tile_rows[(size_t)ntile * kTilePairs + tid + 512];
      same_src = nchunk == chunk && tiles[ntile].src_base == tiles[tile].src_base;
    }
    if constexpr (pipe) {
      {  // bin wave: sources sa; load bin wave + 8's receivers (and sources)
        RecOps<NK16, kSplit> rn;
        load_rec_ops<kMode>(hi, lo, ia, wave + 8, nsta, nspec, lane_g, rn);
        if (sb_stale) load_src_ops<kMode>(hi, lo, ia, wave + 8, nsta, nspec, lane_g, sb);
        // keep the scheduler from sinking the loads below the MFMAs
        __builtin_amdgcn_sched_barrier(0);
        compute_tile_bin<kMode, NK16, kSplit>(sa.sh, sa.sl, rc.rh, rc.rl, ia, wave, lane_r, lane_g,
                                              stage);
        rc = rn;
      }
      {  // bin wave + 8: sources sb; load the next item's bin wave
        RecOps<NK16, kSplit> rn;
        if (nitem >= 0) {
          load_rec_ops<kMode>(hi, lo, na, wave, nsta, nspec, lane_g, rn);
          if (!same_src) load_src_ops<kMode>(hi, lo, na, wave, nsta, nspec, lane_g, sa);
        }
        __builtin_amdgcn_sched_barrier(0);
        compute_tile_bin<kMode, NK16, kSplit>(sb.sh, sb.sl, rc.rh, rc.rl, ia, wave + 8, lane_r,
                                              lane_g, stage);
        rc = rn;
      }
      sb_stale = !same_src;
    } else if constexpr (kChunked) {
#pragma unroll 1
      for (int bi = 0; bi < kBins; ++bi)
        compute_tile_bin_chunked<kMode, NK16, kSplit>(hi, lo, ia, wave + 8 * bi, nsta, nspec,
                                                      lane_r, lane_g, kchunks, stage);
    } else if constexpr (kHoldSrc) {
      // unpipelined, the sources of both bins held across items sharing
      // them; every operand load of a bin is issued before its first MFMA
      {
        RecOps<NK16, kSplit> r;
        load_rec_ops<kMode>(hi, lo, ia, wave, nsta, nspec, lane_g, r);
        if (sb_stale) load_src_ops<kMode>(hi, lo, ia, wave, nsta, nspec, lane_g, sa);
        __builtin_amdgcn_sched_barrier(0);
        compute_tile_bin<kMode, NK16, kSplit>(sa.sh, sa.sl, r.rh, r.rl, ia, wave, lane_r, lane_g,
                                              stage);
      }
      {
        RecOps<NK16, kSplit> r;
        load_rec_ops<kMode>(hi, lo, ia, wave + 8, nsta, nspec, lane_g, r);
        if (sb_stale) load_src_ops<kMode>(hi, lo, ia, wave + 8, nsta, nspec, lane_g, sb);
        __builtin_amdgcn_sched_barrier(0);
        compute_tile_bin<kMode, NK16, kSplit>(sb.sh, sb.sl, r.rh, r.rl, ia, wave + 8, lane_r,
                                              lane_g, stage);
      }
      sb_stale = !same_src;
    } else {
#pragma unroll 1
      for (int bi = 0; bi < kBins; ++bi) {
        TileOps<NK16, kSplit> o;
        load_tile_ops<kMode>(hi, lo, ia, wave + 8 * bi, nsta, nspec, lane_g, o);
        // every operand load of the bin is issued before the first MFMA, so
        // their latencies overlap instead of being waited for one by one
        // (up to kh 64; deeper operands would not fit the registers)
        if constexpr (NK16 <= 8) __builtin_amdgcn_sched_barrier(0);
        compute_tile_bin<kMode>(o, ia, wave + 8 * bi, lane_r, lane_g, stage);
      }
    }
    __syncthreads();
    // 8 consecutive threads write one pair's 16 bins (128 bytes), two
    // adjacent bins (16 bytes) each.  The swizzle keeps bins 2c and 2c+1 in
    // one 16-byte slot, (2c) ^ (p & 14), in swapped order when p is odd.
    const int b = (tid & 7) * 2;
    const int k = ia.k0 + b;
    if (k < nspec) {
      for (int p = tid >> 3; p < kTilePairs; p += blockDim.x >> 3) {
        const int row = rows[p];
        if (row < 0) continue;
        const float4 v = *(const float4 *)(stage + p * kBinChunk + (b ^ (p & 14)));
        const bool sw = (p & 1) != 0;
        const float2 v0 = sw ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
        const float2 v1 = sw ? make_float2(v.x, v.y) : make_float2(v.z, v.w);
        float2 *dst = xspec + (size_t)row * xpitch + k;
        if (k + 1 < nspec)
          store_stream2(dst, v0, v1);
        else
          store_stream(dst, v0);
      }
    }
    item = nitem;
    tile = ntile;
    chunk = nchunk;
    ia = na;
  }
}

// ---------------------------------------------------------------------------
// Lag synthesis.  Persistent: each 1024-thread workgroup owns one CU's LDS
// and walks rows blockIdx.x, +gridDim.x, ...; the next row is loaded into
// registers while the current row's FFT runs, so HBM reads overlap the LDS
// passes.  Per row:
//   out[m] = C2R(X)[(m - H) mod nfft] * (1/(nfft*dt)),  m = 0..2H
// which equals the reference's (-1)^k shift + C2R + extraction of
// [nfft/2 - H, nfft/2 + H].  The scale is evaluated as the reference does:
// weight = 1.0/(double)((float)nfft * dt), product in double, rounded to float.
// ---------------------------------------------------------------------------
// 1024 threads (16 waves, 128 VGPRs each, 40 of them holding the prefetched
// next row).  The butterfly set is a template parameter: radices up to 8, or
// additionally the register-resident composite 9, which takes one LDS pass
// fewer on 2*3^3*7^3 (radix 14 no longer fits the register budget).
constexpr int kLagThreads = 1024;
constexpr long long kLagLargeChunk = 2048;  // rows per four-step lag call
constexpr int kLagMaxRadixDefault = 9;
constexpr unsigned kRadixLag8 = (1u << 2) | (1u << 3) | (1u << 4) | (1u << 5) | (1u << 6) | (1u << 7) | (1u << 8);
constexpr unsigned kRadixLag9 = kRadixLag8 | (1u << 9);
constexpr int kLagRegs = (kLdsComplexCapacity + kLagThreads - 1) / kLagThreads;

__device__ __forceinline__ void lag_prefetch(float2 (&pre)[kLagRegs], const float2 *__restrict__ X,
                                             int nspec) {
#pragma unroll
  for (int i = 0; i < kLagRegs; ++i) {
    const int k = threadIdx.x + i * kLagThreads;
    if (k < nspec) pre[i] = X[k];
  }
}

template <unsigned kRadixLag>
__global__ void __launch_bounds__(kLagThreads)
lag_kernel(const float2 *__restrict__ xspec, long long xpitch, int nspec, long long nrows,
           const int *__restrict__ row_out, float *__restrict__ cc, int half_cc, double weight,
           const float *__restrict__ task_scale, FftPlanDev plan, const float2 *__restrict__ tw,
           const float2 *__restrict__ fold) {
  extern __shared__ float2 lds[];
  const int M = plan.n;  // nspec == M + 1
  stage_real_tables(lds, plan, tw, fold);
  const float2 *twl = lds + M + 1;
  const float2 *fl = twl + tw_lds_count(M);
  float2 pre[kLagRegs];
  long long row = blockIdx.x;
  if (row < nrows) lag_prefetch(pre, xspec + row * xpitch, nspec);
  for (; row < nrows; row += gridDim.x) {
#pragma unroll
    for (int i = 0; i < kLagRegs; ++i) {
      const int k = threadIdx.x + i * kLagThreads;
      if (k < nspec) lds[k] = pre[i];
    }
    __syncthreads();
    const long long nxt = row + gridDim.x;
    if (nxt < nrows) lag_prefetch(pre, xspec + nxt * xpitch, nspec);
    // fold pairs (k, M-k) in place: Z[k] = (X[k] + conj X[M-k]) + i w_k (X[k] - conj X[M-k])
    for (int k = threadIdx.x; 2 * k <= M; k += blockDim.x) {
      if (k == 0) {
        const float a = lds[0].x, b = lds[M].x;  // imaginary parts of DC/Nyquist ignored
        lds[0] = make_float2(a + b, a - b);
      } else {
        const float2 a = lds[k];
        const float2 b = lds[M - k];
        const float2 s1 = cadd(a, cconj(b));
        const float2 d1 = cmul(fold_lookup(fl, k), csub(a, cconj(b)));
        const float2 zk = make_float2(s1.x - d1.y, s1.y + d1.x);
        if (2 * k != M) {
          const float2 s2 = cadd(b, cconj(a));
          const float2 d2 = cmul(fold_lookup(fl, M - k), csub(b, cconj(a)));
          lds[M - k] = make_float2(s2.x - d2.y, s2.y + d2.x);
        }
        lds[k] = zk;
      }
    }
    __syncthreads();
    // all passes but the last; the last (largest radix, stride 1) is
    // evaluated only for the bins that hold kept lags
    lds_fft_head<kRadixLag>(lds, 1, plan, twl, 1.f, plan.npass - 1);
    const int out_row = row_out ? row_out[row] : (int)row;
    float *y = cc + (size_t)out_row * (2 * half_cc + 1);
    const double wr = task_scale ? weight * (double)task_scale[out_row] : weight;
    // lag l = 2n + par reads Re (par 0) or Im (par 1) of Z[n mod M]
    const int n_lo = -((half_cc + 1) / 2), n_hi = half_cc / 2;
    for (int i = threadIdx.x; i <= n_hi - n_lo; i += blockDim.x) {
      const int n = n_lo + i;
      const float2 z = last_pass_bin(lds, plan, twl, n < 0 ? n + M : n);
      const int m0 = 2 * n + half_cc;
      if (m0 >= 0) y[m0] = (float)((double)z.x * wr);
      if (m0 + 1 <= 2 * half_cc) y[m0 + 1] = (float)((double)z.y * wr);
    }
    __syncthreads();
  }
}

// Lag synthesis at the production length nfft = 37044 (3600 s windows at
// 0.1 s): the same steps as lag_kernel on a compile-time plan of the
// 18522-point inverse FFT, 6 x 7 x 7 x 7 x 9, whose last (radix-9, stride-1)
// pass is evaluated only for the kept lags.  Pass 0 reads its twiddles from
// a global table; the next row is prefetched only after that pass, because
// vector-memory counters retire in order and an earlier prefetch would make
// every twiddle read wait for it.  Passes 1-3 take their twiddles from LDS.
// The fold table exp(+2*pi*i*k/nfft), k <= M/2, and the 9th roots of unity
// are global.
using LagPlan18522 = FixedPlan<18522, 6, 7, 7, 7, 9>;
static_assert(LagPlan18522::valid(), "18522 = 6*7*7*7*9");

template <class FP>
__global__ void __launch_bounds__(kLagThreads)
lag_kernel_fixed(const float2 *__restrict__ xspec, long long xpitch, long long nrows,
                 const int *__restrict__ row_out, float *__restrict__ cc, int half_cc,
                 double weight, const float *__restrict__ task_scale,
                 const float2 *__restrict__ twg, const float2 *__restrict__ twl_g,
                 const float2 *__restrict__ foldg, const float2 *__restrict__ roots) {
  extern __shared__ float2 lds[];  // [row: M+1][base twiddles of passes 1..P-2]
  constexpr int M = FP::kM;
  constexpr int nspec = M + 1;
  static_assert(nspec <= kLagRegs * kLagThreads, "prefetch registers cover one row");
  float2 *twl = lds + nspec;
  for (int i = threadIdx.x; i < fixed_lds_tw_count<FP>(); i += kLagThreads) twl[i] = twl_g[i];
  float2 pre[kLagRegs];
  long long row = blockIdx.x;
  if (row < nrows) lag_prefetch(pre, xspec + row * xpitch, nspec);
  for (; row < nrows; row += gridDim.x) {
#pragma unroll
    for (int i = 0; i < kLagRegs; ++i) {
      const int k = threadIdx.x + i * kLagThreads;
      if (k < nspec) lds[k] = pre[i];
    }
    __syncthreads();
    // fold pairs (k, M-k): Z[k] = (X[k] + conj X[M-k]) + i w_k (X[k] - conj X[M-k]),
    // w_{M-k} = -conj(w_k)
    for (int k = threadIdx.x; 2 * k <= M; k += kLagThreads) {
      if (k == 0) {
        const float a = lds[0].x, b = lds[M].x;  // imaginary parts of DC/Nyquist ignored
        lds[0] = make_float2(a + b, a - b);
      } else {
        const float2 w = foldg[k];
        const float2 a = lds[k];
        const float2 b = lds[M - k];
        const float2 s1 = cadd(a, cconj(b));
        const float2 d1 = cmul(w, csub(a, cconj(b)));
        const float2 zk = make_float2(s1.x - d1.y, s1.y + d1.x);
        if (2 * k != M) {
          const float2 s2 = cadd(b, cconj(a));
          const float2 d2 = cmul(make_float2(-w.x, w.y), csub(b, cconj(a)));
          lds[M - k] = make_float2(s2.x - d2.y, s2.y + d2.x);
        }
        lds[k] = zk;
      }
    }
    __syncthreads();
    fixed_dif_pass<FP, 0, kLagThreads, false>(lds, twg, nullptr, 1.f);
    const long long nxt = row + gridDim.x;
    if (nxt < nrows) lag_prefetch(pre, xspec + nxt * xpitch, nspec);
    fixed_dif_passes_lds<FP, 1, kLagThreads>(lds, twl, 1.f);
    const int out_row = row_out ? row_out[row] : (int)row;
    float *y = cc + (size_t)out_row * (2 * half_cc + 1);
    const double wr = task_scale ? weight * (double)task_scale[out_row] : weight;
    const int n_lo = -((half_cc + 1) / 2), n_hi = half_cc / 2;
    for (int i = threadIdx.x; i <= n_hi - n_lo; i += kLagThreads) {
      const int n = n_lo + i;
      const float2 z = fixed_last_pass_bin<FP>(lds, roots, n < 0 ? n + M : n);
      const int m0 = 2 * n + half_cc;
      if (m0 >= 0) y[m0] = (float)((double)z.x * wr);
      if (m0 + 1 <= 2 * half_cc) y[m0 + 1] = (float)((double)z.y * wr);
    }
    __syncthreads();
  }
}

// Lag synthesis at nfft 37044 with the fold and the first DIF pass taken in
// registers straight from the row, so the row reaches LDS once, already
// through pass 0 (the store of the raw row, the fold sweep and pass 0's LDS
// round trip of lag_kernel_fixed disappear, with two barriers).
//
// Pass 0 (radix R0, stride S0 = M/R0) butterfly j reads Z[j + S0 m]; the fold
// of bin k needs X[k] and X[M-k], and M - (j + S0 m) = (S0 - j) + S0 (R0-1-m)
// is leg R0-1-m of butterfly j' = S0 - j.  One thread therefore takes the
// pair (j, j'): it loads the R0 legs of j and the R0 mirrored bins, forms the
// 2 R0 folded values, runs both butterflies and writes them to LDS.  j = 0 is
// its own mirror (leg m pairs with leg R0-m, bin 0 with X[M]).  The fold
// factor of bin j + S0 m is w_j c_m with w_j = exp(2 pi i j/2M) and the
// constant c_m = exp(2 pi i m/(2 R0)); w_{j'} = c_1 conj(w_j).  The pass-0
// twiddles are the powers of t_j = exp(2 pi i j/M), with t_{j'} =
// exp(2 pi i/R0) conj(t_j); w_j and t_j of a thread's pairs stay in
// registers for the whole launch.  The next row is loaded into the same
// registers once pass 0 is in LDS, and every table the later passes read
// (base twiddles, the last radix's roots) is in LDS, so no global read waits
// behind that load (vector-memory counters retire in order).
template <class FP>
__device__ __forceinline__ float2 fold_bin(float2 A, float2 B, float2 w) {
  // Z = A + conj(B) + i w (A - conj(B))
  const float2 s = make_float2(A.x + B.x, A.y - B.y);
  const float2 d = cmul(w, make_float2(A.x - B.x, A.y + B.y));
  return make_float2(s.x - d.y, s.y + d.x);
}

template <class FP>
struct LagPass0 {
  static constexpr int M = FP::kM;
  static constexpr int R0 = FP::kRadix[0];
  static constexpr int S0 = FP::stride(0);
  static constexpr int kPairs = (S0 + 1) / 2;  // j = 0 and the pairs (j, S0 - j)
  static constexpr int kPer = (kPairs + kLagThreads - 1) / kLagThreads;
  static_assert(S0 % 2 == 1, "pass-0 butterflies pair up as (j, S0 - j)");
};

// A cross-spectrum value, read once: nontemporal, so the per-row tables the
// lag kernels read stay in L2.
__device__ __forceinline__ float2 xspec_load(const float2 *p) {
  return __builtin_bit_cast(float2, __builtin_nontemporal_load((const unsigned long long *)p));
}

template <class FP>
__device__ __forceinline__ void lag_r0_load(float2 (&a)[LagPass0<FP>::kPer][FP::kRadix[0]],
                                            float2 (&b)[LagPass0<FP>::kPer][FP::kRadix[0]],
                                            const float2 *__restrict__ X) {
  using L = LagPass0<FP>;
#pragma unroll
  for (int it = 0; it < L::kPer; ++it) {
    const int j = threadIdx.x + it * kLagThreads;
    if (j < L::kPairs) {
      if (j == 0) {
#pragma unroll
        for (int m = 0; m < L::R0; ++m) a[it][m] = xspec_load(X + L::S0 * m);
        b[it][0] = xspec_load(X + L::M);
      } else {
#pragma unroll
        for (int m = 0; m < L::R0; ++m) {
          a[it][m] = xspec_load(X + j + L::S0 * m);
          b[it][m] = xspec_load(X + L::M - j - L::S0 * m);
        }
      }
    }
  }
}

// lag_r0_load of the one-window stack X[k] = conj(Ss[k]) Sr[k], formed from
// the two stations' spectra as it is read.
template <class FP>
__device__ __forceinline__ void lag_r0_load_prod(float2 (&a)[LagPass0<FP>::kPer][FP::kRadix[0]],
                                                 float2 (&b)[LagPass0<FP>::kPer][FP::kRadix[0]],
                                                 const float2 *__restrict__ Ss,
                                                 const float2 *__restrict__ Sr) {
  using L = LagPass0<FP>;
  auto x = [&](int k) {
    const float2 u = Ss[k], v = Sr[k];
    return make_float2(u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x);
  };
#pragma unroll
  for (int it = 0; it < L::kPer; ++it) {
    const int j = threadIdx.x + it * kLagThreads;
    if (j < L::kPairs) {
      if (j == 0) {
#pragma unroll
        for (int m = 0; m < L::R0; ++m) a[it][m] = x(L::S0 * m);
        b[it][0] = x(L::M);
      } else {
#pragma unroll
        for (int m = 0; m < L::R0; ++m) {
          a[it][m] = x(j + L::S0 * m);
          b[it][m] = x(L::M - j - L::S0 * m);
        }
      }
    }
    // one group's loads in flight at a time: both spectra of every group at
    // once would not fit beside the rest of the kernel's registers
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <class FP>
__device__ __forceinline__ void lag_r0_butterfly(float2 *lds, float2 (&u)[FP::kRadix[0]], int j,
                                                 float2 t) {
  using L = LagPass0<FP>;
  dft<L::R0>(u, 1.f);
  float2 tq = t;
#pragma unroll
  for (int q = 1; q < L::R0; ++q) {
    u[q] = cmul(u[q], tq);
    if (q + 1 < L::R0) tq = cmul(tq, t);
  }
#pragma unroll
  for (int q = 0; q < L::R0; ++q) lds[j + L::S0 * q] = u[q];
}

// [Z: M][base twiddles of passes 1..P-2][roots of the last radix, padded to
// an even count][first lag bin of each thread: kLagThreads ints]
template <class FP>
constexpr int lag_r0_lds_count() {
  return FP::kM + fixed_lds_tw_count<FP>() + (FP::kRadix[FP::kPasses - 1] + 1) / 2 * 2 +
         kLagThreads / 2;
}

// kEdge: pass P-2 forms only the outputs the kept lags read
// (fixed_dif_pass_edge; the host checks that half_cc allows it).
// kProd (one window): row r's stack is conj(S[row_src[r]]) S[row_rec[r]],
// formed from the spectra (rows of xpitch) as they are read, so no
// cross-spectrum is written or read; the loads of a row are then issued at
// its start instead of during the previous row.
template <class FP, bool kEdge, bool kProd = false>
__global__ void __launch_bounds__(kLagThreads)
lag_kernel_fixed_r0(const float2 *__restrict__ xspec, long long xpitch, long long nrows,
                    const int *__restrict__ row_out, float *__restrict__ cc, int half_cc,
                    double weight, const float *__restrict__ task_scale,
                    const float2 *__restrict__ twg, const float2 *__restrict__ twl_g,
                    const float2 *__restrict__ foldg, const float2 *__restrict__ roots,
                    const int *__restrict__ row_src = nullptr,
                    const int *__restrict__ row_rec = nullptr) {
  using L = LagPass0<FP>;
  constexpr int R0 = L::R0, S0 = L::S0;
  constexpr int RL = FP::kRadix[FP::kPasses - 1];
  extern __shared__ float2 lds[];  // layout: lag_r0_lds_count
  float2 *twl = lds + L::M;
  float2 *rts = twl + fixed_lds_tw_count<FP>();
  int *firsts = (int *)(rts + (RL + 1) / 2 * 2);
  for (int i = threadIdx.x; i < fixed_lds_tw_count<FP>(); i += kLagThreads) twl[i] = twl_g[i];
  if (threadIdx.x < RL) rts[threadIdx.x] = roots[threadIdx.x];
  constexpr RootTable<2 * R0> cr{};  // c_m = exp(2 pi i m / (2 R0))
  constexpr RootTable<R0> er{};
  const float2 e1 = make_float2(er.c[1], er.s[1]);  // exp(2 pi i / R0)
  float2 a[L::kPer][R0], b[L::kPer][R0];
  // lag bins: n in [n_lo, n_hi]; a thread's first bin has the same
  // last-pass block and output index in every row, kept in LDS as
  // base * 16 + t (a register would spill at this kernel's 128)
  static_assert(RL <= 16, "output index fits 4 bits");
  const int n_lo = -((half_cc + 1) / 2), n_hi = half_cc / 2;
  if (threadIdx.x <= n_hi - n_lo) {
    const int n = n_lo + (int)threadIdx.x;
    int t, base;
    fixed_last_pass_index<FP>(n < 0 ? n + L::M : n, t, base);
    firsts[threadIdx.x] = base * 16 + t;
  }
  long long row = blockIdx.x;
  if constexpr (!kProd)
    if (row < nrows) lag_r0_load<FP>(a, b, xspec + row * xpitch);
  for (; row < nrows; row += gridDim.x) {
    if constexpr (kProd)
      lag_r0_load_prod<FP>(a, b, xspec + (size_t)row_src[row] * xpitch,
                           xspec + (size_t)row_rec[row] * xpitch);
#pragma unroll
    for (int it = 0; it < L::kPer; ++it) {
      const int j = threadIdx.x + it * kLagThreads;
      if (j >= L::kPairs) continue;
      // bases of the pair, read per row (L2): waiting for them waits for the
      // row's own loads, which this pass needs anyway
      const float2 wjv = foldg[j];
      const float2 tjv = twg[j * FP::tw_pitch(0)];
      float2 u[R0];
      if (j == 0) {
        u[0] = make_float2(a[it][0].x + b[it][0].x, a[it][0].x - b[it][0].x);
#pragma unroll
        for (int m = 1; m < R0; ++m)
          u[m] = fold_bin<FP>(a[it][m], a[it][R0 - m], make_float2(cr.c[m], cr.s[m]));
        lag_r0_butterfly<FP>(lds, u, 0, make_float2(1.f, 0.f));
      } else {
        float2 v[R0];
#pragma unroll
        for (int m = 0; m < R0; ++m) {
          // the bin and its mirror share s = A + conj B and p = w (A - conj B):
          //   Z = s + i p,  Z' = B + conj A + i (-conj w)(B - conj A) = conj s + i conj p
          const float2 w = cmul(wjv, make_float2(cr.c[m], cr.s[m]));
          const float2 A = a[it][m], B = b[it][m];
          const float2 sv = make_float2(A.x + B.x, A.y - B.y);
          const float2 pv = cmul(w, make_float2(A.x - B.x, A.y + B.y));
          u[m] = make_float2(sv.x - pv.y, sv.y + pv.x);
          v[R0 - 1 - m] = make_float2(sv.x + pv.y, pv.x - sv.y);
        }
        lag_r0_butterfly<FP>(lds, u, j, tjv);
        lag_r0_butterfly<FP>(lds, v, S0 - j, cmul(e1, cconj(tjv)));
      }
    }
    __syncthreads();
    const long long nxt = row + gridDim.x;
    if constexpr (!kProd)
      if (nxt < nrows) lag_r0_load<FP>(a, b, xspec + nxt * xpitch);
    if constexpr (kEdge) {
      fixed_dif_passes_lds<FP, 1, kLagThreads, FP::kPasses - 2>(lds, twl, 1.f);
      fixed_dif_pass_edge<FP, kLagThreads>(lds, twl, 1.f);
    } else {
      fixed_dif_passes_lds<FP, 1, kLagThreads>(lds, twl, 1.f);
    }
    const int out_row = row_out ? row_out[row] : (int)row;
    float *y = cc + (size_t)out_row * (2 * half_cc + 1);
    const double wr = task_scale ? weight * (double)task_scale[out_row] : weight;
    for (int i = threadIdx.x; i <= n_hi - n_lo; i += kLagThreads) {
      const int n = n_lo + i;
      float2 z;
      if (i == (int)threadIdx.x) {
        const int first = firsts[threadIdx.x];
        z = fixed_last_pass_at<FP, true>(lds, rts, first & 15, first >> 4);
      } else {
        z = fixed_last_pass_bin<FP, true>(lds, rts, n < 0 ? n + L::M : n);
      }
      const int m0 = 2 * n + half_cc;
      if (m0 >= 0) y[m0] = (float)((double)z.x * wr);
      if (m0 + 1 <= 2 * half_cc) y[m0 + 1] = (float)((double)z.y * wr);
    }
    __syncthreads();
  }
}

// PLAIN form, one launch per window w (the reference's accumulation order):
//   stack[r][k] += (conj(S[w][src_r][k]) * S[w][rec_r][k]) * scale
// Work items are (row, 256-bin block) pairs, nrows * nbk of them, visited
// with a grid-stride loop so the launch stays within the 2^32 work-item
// limit of one dispatch.
__global__ void __launch_bounds__(256)
plain_accum_kernel(const float2 *__restrict__ spec_w, int nspec, const int *__restrict__ row_src,
                   const int *__restrict__ row_rec, long long nrows, float2 *__restrict__ stack,
                   long long xpitch, int nbk, float scale) {
  const long long nitems = nrows * nbk;
  for (long long it = blockIdx.x; it < nitems; it += gridDim.x) {
    const long long row = it / nbk;
    const int k = (int)(it % nbk) * 256 + threadIdx.x;
    if (k >= nspec) continue;
    const float2 a = spec_w[(size_t)row_src[row] * nspec + k];
    const float2 b = spec_w[(size_t)row_rec[row] * nspec + k];
    const float pr = a.x * b.x + a.y * b.y;
    const float pi = a.x * b.y - a.y * b.x;
    float2 *d = stack + (size_t)row * xpitch + k;
    float2 v = *d;
    v.x += pr * scale;
    v.y += pi * scale;
    *d = v;
  }
}

// Lags of rows whose folded inverse transform z (M = nfft/2 complex per row,
// z[j] = y[2j] + i y[2j+1]) is done: out[m] = y[(m - H) mod nfft] * weight,
// the product in double (the reference's (-1)^k shift + C2R + extraction).
__global__ void lag_extract_kernel(const float2 *__restrict__ z, int M, long long nrows,
                                   const int *__restrict__ row_out, float *__restrict__ cc,
                                   int half_cc, double weight,
                                   const float *__restrict__ task_scale) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncc = 2 * half_cc + 1;
  if (m >= ncc) return;
  for (long long r = blockIdx.y; r < nrows; r += gridDim.y) {
    const int out_row = row_out ? row_out[r] : (int)r;
    const double wr = task_scale ? weight * (double)task_scale[out_row] : weight;
    int i = m - half_cc;
    if (i < 0) i += 2 * M;
    const float2 v = z[r * M + (i >> 1)];
    cc[(size_t)out_row * ncc + m] = (float)((double)((i & 1) ? v.y : v.x) * wr);
  }
}

// Copies workspace rows to their task rows (cross-spectra output order).
__global__ void scatter_rows_kernel(const float2 *__restrict__ src, long long src_pitch,
                                    const int *__restrict__ row_task, int nspec,
                                    float2 *__restrict__ dst) {
  const size_t r = blockIdx.x;
  const size_t t = (size_t)row_task[r];
  for (int k = threadIdx.x; k < nspec; k += blockDim.x)
    dst[t * nspec + k] = src[r * src_pitch + k];
}

}  // namespace fxc

using namespace fxc;

// Operand depth kh of the MFMA forms for nstep windows (K' = 2*kh).  The
// contraction's time follows K' (operand loads), so kh pads as little as
// the compiled depths allow: 8 and 16, then multiples of 8 up to 48, then
// 64, 96 and 128 (depths 80 and 112 would spill registers).  At depths
// that are not multiples of 32 the receiver's [Im|-Re] operand is loaded
// rather than formed in registers, which costs far less than the padding
// (47 windows, 1000 stations: contraction 29.9 ms at kh 48, 45.3 ms at 64).
static int xc_operand_depth(int nstep) {
  if (nstep <= 48) return nstep <= 8 ? 8 : (nstep + 7) / 8 * 8;
  if (nstep <= 128) return nstep <= 64 ? 64 : nstep <= 96 ? 96 : 128;
  return (nstep + 31) / 32 * 32;  // chunks of 32 windows (see xc_depth_chunks)
}

// Chunks of 32 windows the contraction runs over for an operand depth kh
// beyond the compiled depths (1 up to kh 128).
static int xc_depth_chunks(int kh) { return kh > 128 ? kh / 32 : 1; }

// ---------------------------------------------------------------------------
// Plan object behind the C ABI.
// ---------------------------------------------------------------------------
struct fxc_xc_plan {
  int nspec = 0, nstep = 0, nfft = 0, half_cc = 0, cc_size = 0;
  long long xpitch = 0;  // workspace row pitch (complex), a multiple of 16 (128 bytes)
  float dt = 0.f;
  int precision = FXC_PREC_FP32;
  int kh = 0;
  int lag_max_radix = 8;           // largest butterfly of the lag FFT plan
  // tables of the compile-time lag plan (nfft/2 == LagPlan18522::kM only)
  float2 *d_fix_tw = nullptr, *d_fix_twl = nullptr, *d_fix_fold = nullptr, *d_fix_roots = nullptr;
  long long chunk_rows_max = 0;  // cross-spectrum rows per workspace fill
  // spec mode (fxc_xc_set_band): the stack is formed for the 16-bin chunks
  // [band_c0, band_c1) only; the workspace columns outside them hold zeros
  // for the first band_zero_rows rows
  int band_c0 = 0, band_c1 = -1;
  size_t band_zero_rows = 0;
  // nfft/2 beyond one workgroup's LDS: lag synthesis by the four-step
  // inverse transform (work buffer for kLagLargeChunk rows)
  bool large_lag = false;
  float2 *d_lwork = nullptr;
  // pairs
  int nsta = 0;
  long long npairs = 0;
  std::vector<TileDesc> h_tiles;
  std::vector<long long> chunk_tile_begin;  // tile index where each chunk starts
  std::vector<long long> chunk_row_begin;   // first workspace row id of each chunk
  TileDesc *d_tiles = nullptr;
  int *d_tile_rows = nullptr;     // chunk-relative rows
  int *d_row_task = nullptr;      // global row -> task index
  int *d_row_src = nullptr;       // global row -> src station (PLAIN form)
  int *d_row_rec = nullptr;       // global row -> rec station (PLAIN form)
  // grow-only capacities of the device index arrays, and the host copies
  // the uploads read from (kept until the next set_pairs, so an upload
  // ordered on a stream never reads freed memory)
  size_t cap_tiles = 0, cap_tile_rows = 0, cap_rows = 0, cap_row_src = 0, cap_row_rec = 0;
  std::vector<int> h_tile_rows, h_row_task, h_row_src, h_row_rec;
  // workspaces; ws_shared: d_xspec, d_hi and d_lo belong to a batch that
  // runs its plans one after another on one stream (not freed or grown here)
  float2 *d_xspec = nullptr;
  size_t xspec_rows = 0;
  unsigned short *d_hi = nullptr, *d_lo = nullptr;
  size_t split_elems = 0;
  bool ws_shared = false;
  unsigned int *d_amax = nullptr;
  float *d_inv_scale = nullptr, *d_scale = nullptr;
  int scale_cap = 0;
  // optional per-kernel timing with events on the launch stream
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;  // (kind, start/stop)
  double t_ms[3] = {0, 0, 0};  // split, contract, lag
  long long t_launch[3] = {0, 0, 0};
};

enum { kTimeSplit = 0, kTimeContract = 1, kTimeLag = 2 };

static int xc_mark(fxc_xc_plan *p, int kind, bool start, hipStream_t st) {
  if (!p->timing) return FXC_OK;
  if (start) {
    hipEvent_t a, b;
    FXC_HIP(hipEventCreate(&a));
    FXC_HIP(hipEventCreate(&b));
    p->events.push_back({kind, {a, b}});
    FXC_HIP(hipEventRecord(a, st));
  } else {
    FXC_HIP(hipEventRecord(p->events.back().second.second, st));
  }
  return FXC_OK;
}

static void xc_free_pairs(fxc_xc_plan *p) {
  dfree(p->d_tiles);
  dfree(p->d_tile_rows);
  dfree(p->d_row_task);
  dfree(p->d_row_src);
  dfree(p->d_row_rec);
  p->d_tiles = nullptr;
  p->d_tile_rows = nullptr;
  p->d_row_task = nullptr;
  p->d_row_src = nullptr;
  p->d_row_rec = nullptr;
  p->cap_tiles = p->cap_tile_rows = p->cap_rows = p->cap_row_src = p->cap_row_rec = 0;
  p->h_tiles.clear();
  p->chunk_tile_begin.clear();
  p->chunk_row_begin.clear();
  p->npairs = 0;
}

// Grows a device array to at least n elements (contents not kept).
template <class T>
static int xc_grow(T **d, size_t *cap, size_t n) {
  if (n <= *cap && *d) return FXC_OK;
  dfree(*d);
  *d = nullptr;
  *cap = 0;
  FXC_TRY(dmalloc((void **)d, std::max<size_t>(n, 1) * sizeof(T)));
  *cap = std::max<size_t>(n, 1);
  return FXC_OK;
}

extern "C" int fxc_xc_plan_create(fxc_xc_plan **out, int nspec, int nstep, float dt, int half_cc,
                                  int precision) {
  FXC_CHECK_ARG(out != nullptr, "fxc_xc_plan_create: null output");
  *out = nullptr;
  FXC_CHECK_ARG(nspec >= 2 && nstep >= 1 && dt > 0.f && half_cc >= 0,
                "fxc_xc_plan_create: invalid shape nspec=%d nstep=%d dt=%g half_cc=%d", nspec,
                nstep, (double)dt, half_cc);
  FXC_CHECK_ARG(precision == FXC_PREC_FP32 || precision == FXC_PREC_BF16X3 ||
                    precision == FXC_PREC_FP16 || precision == FXC_PREC_PLAIN ||
                    precision == FXC_PREC_FP16X3,
                "fxc_xc_plan_create: unknown precision %d", precision);
  const int nfft = 2 * (nspec - 1);
  FXC_CHECK_ARG(2 * half_cc + 1 <= nfft, "fxc_xc_plan_create: 2*half_cc+1 exceeds nfft");
  // lag butterfly set: radices up to 9 (default) or 8; FXC_XC_LAG_MAX_RADIX overrides
  int lag_max_radix = kLagMaxRadixDefault;
  if (const char *env = getenv("FXC_XC_LAG_MAX_RADIX")) lag_max_radix = atoi(env) >= 9 ? 9 : 8;
  const bool large_lag = real_lds_count(nfft / 2) > kLdsComplexCapacity;
  if (large_lag) {
    int a = 0, b = 0;
    if (!large_factor(nfft / 2, &a, &b)) {
      set_error("fxc_xc_plan_create: nfft %d exceeds one workgroup and has no two-level "
                "factorisation", nfft);
      return FXC_ERR_UNSUPPORTED;
    }
  } else {
    const RealFftPlan *rp = nullptr;
    FXC_TRY(plan_cache_real(nfft, 1, &rp, lag_max_radix, true));
  }
  fxc_xc_plan *p = new fxc_xc_plan();
  p->large_lag = large_lag;
  p->nspec = nspec;
  p->nstep = nstep;
  p->nfft = nfft;
  p->half_cc = half_cc;
  p->cc_size = 2 * half_cc + 1;
  p->dt = dt;
  p->precision = precision;
  p->xpitch = (nspec + 15) / 16 * 16;
  p->kh = xc_operand_depth(nstep);
  p->chunk_rows_max = 0;
  p->lag_max_radix = lag_max_radix;
  const char *fixed_env = getenv("FXC_XC_LAG_FIXED");
  if (nfft / 2 == LagPlan18522::kM && !(fixed_env && atoi(fixed_env) == 0)) {
    using FP = LagPlan18522;
    const std::vector<float2> tw = fixed_twiddles<FP>(1, 1);
    const std::vector<float2> twl = fixed_lds_twiddles<FP>(1);
    std::vector<float2> fold(FP::kM / 2 + 1), roots(FP::kRadix[FP::kPasses - 1]);
    for (size_t k = 0; k < fold.size(); ++k) {
      const double a = 2.0 * M_PI * (double)k / (double)nfft;
      fold[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    for (size_t u = 0; u < roots.size(); ++u) {
      const double a = 2.0 * M_PI * (double)u / (double)roots.size();
      roots[u] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    auto upload = [](float2 **dst, const std::vector<float2> &v) -> int {
      FXC_TRY(dmalloc((void **)dst, v.size() * sizeof(float2)));
      FXC_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(float2), hipMemcpyHostToDevice));
      return FXC_OK;
    };
    int rc = upload(&p->d_fix_tw, tw);
    if (rc == FXC_OK) rc = upload(&p->d_fix_twl, twl);
    if (rc == FXC_OK) rc = upload(&p->d_fix_fold, fold);
    if (rc == FXC_OK) rc = upload(&p->d_fix_roots, roots);
    if (rc != FXC_OK) {
      fxc_xc_plan_destroy(p);
      return rc;
    }
  }
  *out = p;
  return FXC_OK;
}

static void xc_drop_events(fxc_xc_plan *p) {
  for (auto &e : p->events) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  p->events.clear();
}

// Enables (1) or disables (0) per-kernel event timing and clears the totals.
extern "C" int fxc_xc_set_timing(fxc_xc_plan *p, int enable) {
  FXC_CHECK_ARG(p, "fxc_xc_set_timing: null plan");
  xc_drop_events(p);
  p->timing = enable != 0;
  for (int i = 0; i < 3; ++i) {
    p->t_ms[i] = 0;
    p->t_launch[i] = 0;
  }
  return FXC_OK;
}

// Waits for the recorded events and returns accumulated kernel milliseconds
// (split, contraction, lag synthesis) and launch counts since the last reset.
extern "C" int fxc_xc_get_timing(fxc_xc_plan *p, double *ms3, long long *launches3) {
  FXC_CHECK_ARG(p && ms3, "fxc_xc_get_timing: invalid arguments");
  for (auto &e : p->events) {
    FXC_HIP(hipEventSynchronize(e.second.second));
    float ms = 0.f;
    FXC_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    p->t_ms[e.first] += ms;
    p->t_launch[e.first] += 1;
  }
  xc_drop_events(p);
  for (int i = 0; i < 3; ++i) {
    ms3[i] = p->t_ms[i];
    if (launches3) launches3[i] = p->t_launch[i];
  }
  return FXC_OK;
}

extern "C" int fxc_xc_plan_destroy(fxc_xc_plan *p) {
  if (!p) return FXC_OK;
  xc_drop_events(p);
  xc_free_pairs(p);
  if (!p->ws_shared) {
    dfree(p->d_xspec);
    dfree(p->d_hi);
    dfree(p->d_lo);
  }
  dfree(p->d_amax);
  dfree(p->d_inv_scale);
  dfree(p->d_scale);
  dfree(p->d_fix_tw);
  dfree(p->d_fix_twl);
  dfree(p->d_fix_fold);
  dfree(p->d_fix_roots);
  dfree(p->d_lwork);
  delete p;
  return FXC_OK;
}

extern "C" int fxc_xc_set_chunk_rows(fxc_xc_plan *p, long long rows) {
  FXC_CHECK_ARG(p && rows >= 0, "fxc_xc_set_chunk_rows: invalid arguments");
  p->chunk_rows_max = rows;
  return FXC_OK;
}

// Builds the 32x32 station tiles for the task list (src[t], rec[t]) and the
// row maps.  Tasks keep their index as the output row of fxc_xc_correlate.
// Task list of a plan.  With `stream` set the index uploads are ordered on
// it and the device arrays only grow, so a caller that reuses a plan with at
// most the pairs it has seen neither allocates nor synchronises the device.
int xc_set_pairs_on(fxc_xc_plan *p, int nsta, const int *h_src, const int *h_rec, long long npairs,
                    hipStream_t stream) {
  FXC_CHECK_ARG(p && nsta >= 1 && npairs >= 0 && (npairs == 0 || (h_src && h_rec)),
                "fxc_xc_set_pairs: invalid arguments");
  p->h_tiles.clear();
  p->chunk_tile_begin.clear();
  p->chunk_row_begin.clear();
  p->npairs = 0;
  p->nsta = nsta;
  if (npairs == 0) return FXC_OK;
  const int nblk = (nsta + kTile - 1) / kTile;
  // tile key -> list of (local pair index, task)
  std::map<long long, std::vector<std::pair<int, long long>>> by_tile;
  for (long long t = 0; t < npairs; ++t) {
    const int s = h_src[t], r = h_rec[t];
    FXC_CHECK_ARG(s >= 0 && s < nsta && r >= 0 && r < nsta,
                  "fxc_xc_set_pairs: pair %lld station index out of range", t);
    const long long key = (long long)(s / kTile) * nblk + (r / kTile);
    by_tile[key].push_back({(s % kTile) * kTile + (r % kTile), t});
  }
  const long long cap = p->chunk_rows_max > 0 ? p->chunk_rows_max : npairs;
  std::vector<int> &rows = p->h_tile_rows;
  std::vector<int> &row_task = p->h_row_task;
  rows.clear();
  row_task.clear();
  rows.reserve(by_tile.size() * kTilePairs);
  row_task.reserve(npairs);
  long long chunk_rows = 0;
  for (auto &kv : by_tile) {
    auto &lst = kv.second;
    std::sort(lst.begin(), lst.end());
    // a local pair listed twice (duplicate task) keeps its first row; later
    // duplicates are computed by their own tile copy
    std::vector<std::vector<std::pair<int, long long>>> copies(1);
    for (auto &e : lst) {
      size_t c = 0;
      while (c < copies.size() && !copies[c].empty() && copies[c].back().first == e.first) ++c;
      if (c == copies.size()) copies.emplace_back();
      copies[c].push_back(e);
    }
    for (auto &cp : copies) {
      if (cp.empty()) continue;
      if (p->h_tiles.empty() || chunk_rows + (long long)cp.size() > cap) {
        p->chunk_tile_begin.push_back((long long)p->h_tiles.size());
        p->chunk_row_begin.push_back((long long)row_task.size());
        chunk_rows = 0;
      }
      TileDesc td;
      td.src_base = (int)(kv.first / nblk) * kTile;
      td.rec_base = (int)(kv.first % nblk) * kTile;
      p->h_tiles.push_back(td);
      const size_t off = rows.size();
      rows.resize(off + kTilePairs, -1);
      for (auto &e : cp) {
        rows[off + e.first] = (int)chunk_rows++;
        row_task.push_back((int)e.second);
      }
    }
  }
  p->chunk_tile_begin.push_back((long long)p->h_tiles.size());
  p->chunk_row_begin.push_back((long long)row_task.size());
  FXC_TRY(xc_grow(&p->d_tiles, &p->cap_tiles, p->h_tiles.size()));
  FXC_TRY(xc_grow(&p->d_tile_rows, &p->cap_tile_rows, rows.size()));
  FXC_TRY(xc_grow(&p->d_row_task, &p->cap_rows, row_task.size()));
  auto up = [&](void *dst, const void *src, size_t bytes) -> int {
    if (stream)
      FXC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    else
      FXC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FXC_OK;
  };
  FXC_TRY(up(p->d_tiles, p->h_tiles.data(), p->h_tiles.size() * sizeof(TileDesc)));
  FXC_TRY(up(p->d_tile_rows, rows.data(), rows.size() * sizeof(int)));
  FXC_TRY(up(p->d_row_task, row_task.data(), row_task.size() * sizeof(int)));
  if (p->precision == FXC_PREC_PLAIN || p->nstep == 1) {
    std::vector<int> &rs = p->h_row_src, &rr = p->h_row_rec;
    rs.resize(row_task.size());
    rr.resize(row_task.size());
    for (size_t i = 0; i < row_task.size(); ++i) {
      rs[i] = h_src[row_task[i]];
      rr[i] = h_rec[row_task[i]];
    }
    FXC_TRY(xc_grow(&p->d_row_src, &p->cap_row_src, rs.size()));
    FXC_TRY(xc_grow(&p->d_row_rec, &p->cap_row_rec, rr.size()));
    FXC_TRY(up(p->d_row_src, rs.data(), rs.size() * sizeof(int)));
    FXC_TRY(up(p->d_row_rec, rr.data(), rr.size() * sizeof(int)));
  }
  p->npairs = npairs;  // only a complete task list is ever used
  return FXC_OK;
}

extern "C" int fxc_xc_set_pairs(fxc_xc_plan *p, int nsta, const int *h_src, const int *h_rec,
                                long long npairs) {
  return xc_set_pairs_on(p, nsta, h_src, h_rec, npairs, nullptr);
}

extern "C" int fxc_xc_pairs_info(const fxc_xc_plan *p, long long *ntiles, long long *nchunks,
                                 long long *max_chunk_rows) {
  FXC_CHECK_ARG(p, "fxc_xc_pairs_info: null plan");
  const long long nc = p->chunk_tile_begin.empty() ? 0 : (long long)p->chunk_tile_begin.size() - 1;
  long long mx = 0;
  for (long long c = 0; c < nc; ++c)
    mx = std::max(mx, p->chunk_row_begin[c + 1] - p->chunk_row_begin[c]);
  if (ntiles) *ntiles = (long long)p->h_tiles.size();
  if (nchunks) *nchunks = nc;
  if (max_chunk_rows) *max_chunk_rows = mx;
  return FXC_OK;
}

// Spec mode: bins [0, k0) and [k1, nspec) of every workspace row set to 0
// (the band's contraction never writes them).
__global__ void band_zero_kernel(float2 *__restrict__ x, long long pitch, long long rows, int k0,
                                 int k1, int nspec) {
  const int nout = k0 + (nspec - k1);
  for (long long r = blockIdx.y; r < rows; r += gridDim.y)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nout; i += gridDim.x * blockDim.x)
      x[r * pitch + (i < k0 ? i : k1 + (i - k0))] = make_float2(0.f, 0.f);
}

static int xc_ensure_workspace(fxc_xc_plan *p, size_t rows) {
  if (p->ws_shared) {
    FXC_CHECK_ARG(rows <= p->xspec_rows &&
                      (size_t)p->nsta * p->nspec * (size_t)(2 * p->kh) <= p->split_elems,
                  "xc: shared workspace too small");
    return FXC_OK;
  }
  if (rows > p->xspec_rows) {
    dfree(p->d_xspec);
    p->d_xspec = nullptr;
    p->xspec_rows = 0;
    FXC_TRY(dmalloc((void **)&p->d_xspec, rows * (size_t)p->xpitch * sizeof(float2)));
    p->xspec_rows = rows;
    p->band_zero_rows = 0;
  }
  if (p->precision == FXC_PREC_BF16X3 || p->precision == FXC_PREC_FP16 ||
      p->precision == FXC_PREC_FP16X3) {
    const size_t elems = (size_t)p->nsta * p->nspec * (size_t)(2 * p->kh);
    if (elems > p->split_elems) {
      dfree(p->d_hi);
      dfree(p->d_lo);
      p->d_hi = p->d_lo = nullptr;
      p->split_elems = 0;
      FXC_TRY(dmalloc((void **)&p->d_hi, elems * sizeof(unsigned short)));
      if (p->precision == FXC_PREC_BF16X3 || p->precision == FXC_PREC_FP16X3)
        FXC_TRY(dmalloc((void **)&p->d_lo, elems * sizeof(unsigned short)));
      p->split_elems = elems;
    }
    if (p->nsta > p->scale_cap) {
      dfree(p->d_amax);
      dfree(p->d_inv_scale);
      dfree(p->d_scale);
      FXC_TRY(dmalloc((void **)&p->d_amax, p->nsta * sizeof(unsigned int)));
      FXC_TRY(dmalloc((void **)&p->d_inv_scale, p->nsta * sizeof(float)));
      FXC_TRY(dmalloc((void **)&p->d_scale, p->nsta * sizeof(float)));
      p->scale_cap = p->nsta;
    }
  }
  return FXC_OK;
}

// Spec mode: zero the out-of-band columns of the workspace rows not yet cleared.
static int xc_band_prepare(fxc_xc_plan *p, hipStream_t st) {
  if (p->band_c1 < 0 || p->band_zero_rows >= p->xspec_rows) return FXC_OK;
  const int k0 = p->band_c0 * kBinChunk, k1 = std::min(p->nspec, p->band_c1 * kBinChunk);
  if (k0 + (p->nspec - k1) > 0) {
    const long long rows = (long long)p->xspec_rows;
    hipLaunchKernelGGL(band_zero_kernel, dim3(16, (unsigned)std::min<long long>(rows, 65535)),
                       dim3(256), 0, st, p->d_xspec, p->xpitch, rows, k0, k1, p->nspec);
    FXC_HIP(hipGetLastError());
  }
  p->band_zero_rows = p->xspec_rows;
  return FXC_OK;
}

// Split pass for the MFMA forms (no-op for FP32).
static int xc_prepare_spectra(fxc_xc_plan *p, const float2 *d_spec, hipStream_t st) {
  if (p->precision == FXC_PREC_FP32 || p->precision == FXC_PREC_PLAIN) return FXC_OK;
  FXC_TRY(xc_mark(p, kTimeSplit, true, st));
  // spec mode: only the band's chunks (kSplitBins == kBinChunk)
  static_assert(kSplitBins == kBinChunk, "split blocks are contraction chunks");
  const bool band = p->band_c1 >= 0;
  const int bin0 = band ? p->band_c0 * kSplitBins : 0;
  const dim3 grid((unsigned)(band ? p->band_c1 - p->band_c0 : ceil_div(p->nspec, kSplitBins)),
                  (unsigned)ceil_div(p->nsta, kSplitSta));
  if (p->precision == FXC_PREC_FP16 || p->precision == FXC_PREC_FP16X3) {
    FXC_HIP(hipMemsetAsync(p->d_amax, 0, p->nsta * sizeof(unsigned int), st));
    hipLaunchKernelGGL(station_absmax_kernel, dim3(8, p->nsta), dim3(256), 0, st, d_spec,
                       p->nstep, p->nsta, p->nspec, p->d_amax);
    hipLaunchKernelGGL(station_scale_kernel, dim3((unsigned)ceil_div(p->nsta, 256)), dim3(256), 0,
                       st, p->d_amax, p->nsta, p->d_inv_scale, p->d_scale);
    if (p->precision == FXC_PREC_FP16)
      hipLaunchKernelGGL(split_kernel<FXC_PREC_FP16>, grid, dim3(kSplitThreads), 0, st, d_spec, p->nstep,
                         p->nsta, p->nspec, p->kh, p->d_inv_scale, p->d_hi, p->d_lo, bin0);
    else
      hipLaunchKernelGGL(split_kernel<FXC_PREC_FP16X3>, grid, dim3(kSplitThreads), 0, st, d_spec, p->nstep,
                         p->nsta, p->nspec, p->kh, p->d_inv_scale, p->d_hi, p->d_lo, bin0);
  } else {
    hipLaunchKernelGGL(split_kernel<FXC_PREC_BF16X3>, grid, dim3(kSplitThreads), 0, st, d_spec, p->nstep,
                       p->nsta, p->nspec, p->kh, (const float *)nullptr, p->d_hi, p->d_lo, bin0);
  }
  FXC_HIP(hipGetLastError());
  FXC_TRY(xc_mark(p, kTimeSplit, false, st));
  return FXC_OK;
}

template <int kMode, int KH, bool kChunked = false>
static int launch_mfma_kh(dim3 grid, size_t lds, hipStream_t st, const unsigned short *hi,
                          const unsigned short *lo, int nsta, int nspec, const float *sc,
                          const TileDesc *tiles, const int *trows, float2 *dst, long long xpitch,
                          float scale, int kchunks, int chunk0, int nchunks) {
  auto kern = contract_mfma_kernel<kMode, KH / 8, kChunked>;
  // persistent grid: a multiple of 8 workgroups (one per CU, LDS-bound),
  // capped by the number of work items
  int dev = 0, ncu = 0;
  FXC_HIP(hipGetDevice(&dev));
  FXC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  unsigned g = (unsigned)std::max(8, ncu / 8 * 8);
  if (const char *env = getenv("FXC_XC_MFMA_GRID")) g = (unsigned)std::max(8, atoi(env) / 8 * 8);
  const dim3 grid1(std::min(g, grid.x));
  // the kernel's item order gives each XCD (blockIdx.x mod 8) its own chunks
  // and covers every item only when the grid is a multiple of 8
  if (grid1.x % 8 != 0) {
    set_error("MFMA contraction: grid %u is not a multiple of 8", grid1.x);
    return FXC_ERR_UNSUPPORTED;
  }
  FXC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  hipLaunchKernelGGL(kern, grid1, dim3(512), lds, st, hi, lo, nsta, nspec, sc, tiles, trows, dst,
                     xpitch, (int)(grid.y), scale, kchunks, chunk0, nchunks);
  return FXC_OK;
}

template <int kMode>
static int launch_mfma_mode(int kh, dim3 grid, size_t lds, hipStream_t st,
                            const unsigned short *hi, const unsigned short *lo, int nsta, int nspec,
                            const float *sc, const TileDesc *tiles, const int *trows, float2 *dst,
                            long long xpitch, float scale, int chunk0, int nchunks) {
  if (xc_depth_chunks(kh) > 1)
    return launch_mfma_kh<kMode, 32, true>(grid, lds, st, hi, lo, nsta, nspec, sc, tiles, trows,
                                           dst, xpitch, scale, xc_depth_chunks(kh), chunk0,
                                           nchunks);
#define FXC_KH(V) \
  case V: return launch_mfma_kh<kMode, V>(grid, lds, st, hi, lo, nsta, nspec, sc, tiles, trows, dst, xpitch, scale, 1, chunk0, nchunks);
  switch (kh) {
    FXC_KH(8) FXC_KH(16) FXC_KH(24) FXC_KH(32) FXC_KH(40) FXC_KH(48) FXC_KH(64) FXC_KH(96)
    FXC_KH(128)
    default: break;
  }
#undef FXC_KH
  set_error("MFMA contraction: unsupported window padding %d", kh);
  return FXC_ERR_UNSUPPORTED;
}

static int launch_mfma_contract(int precision, int kh, dim3 grid, size_t lds, hipStream_t st,
                                const unsigned short *hi, const unsigned short *lo, int nsta,
                                int nspec, const float *sc, const TileDesc *tiles,
                                const int *trows, float2 *dst, long long xpitch, float scale,
                                int chunk0, int nchunks) {
  if (precision == FXC_PREC_FP16X3)
    return launch_mfma_mode<FXC_PREC_FP16X3>(kh, grid, lds, st, hi, lo, nsta, nspec, sc, tiles,
                                             trows, dst, xpitch, scale, chunk0, nchunks);
  if (precision == FXC_PREC_BF16X3)
    return launch_mfma_mode<FXC_PREC_BF16X3>(kh, grid, lds, st, hi, lo, nsta, nspec, sc, tiles,
                                             trows, dst, xpitch, scale, chunk0, nchunks);
  return launch_mfma_mode<FXC_PREC_FP16>(kh, grid, lds, st, hi, lo, nsta, nspec, sc, tiles, trows,
                                         dst, xpitch, scale, chunk0, nchunks);
}

static int xc_contract_chunk(fxc_xc_plan *p, const float2 *d_spec, long long c, float2 *dst,
                             hipStream_t st) {
  const long long t0 = p->chunk_tile_begin[c], t1 = p->chunk_tile_begin[c + 1];
  const unsigned ntl = (unsigned)(t1 - t0);
  int c0 = 0, nbc = ceil_div(p->nspec, kBinChunk);
  if (p->band_c1 >= 0 && p->precision != FXC_PREC_PLAIN) {
    c0 = p->band_c0;
    nbc = p->band_c1 - p->band_c0;
  }
  // 1-D grid covering ntl tiles x nbc chunks in groups of kChunkGroup chunks
  const unsigned nwg = ntl * (unsigned)(ceil_div(nbc, kChunkGroup) * kChunkGroup);
  const float scale = 1.0f / (float)p->nstep;
  const TileDesc *tiles = p->d_tiles + t0;
  const int *trows = p->d_tile_rows + t0 * kTilePairs;
  FXC_TRY(xc_mark(p, kTimeContract, true, st));
  if (p->precision == FXC_PREC_PLAIN) {
    const long long r0 = p->chunk_row_begin[c], nrows = p->chunk_row_begin[c + 1] - r0;
    const int nbk = ceil_div(p->nspec, 256);
    FXC_HIP(hipMemsetAsync(dst, 0, (size_t)nrows * p->xpitch * sizeof(float2), st));
    const size_t win = (size_t)p->nsta * p->nspec;
    const long long grid = std::min<long long>(nrows * nbk, 1ll << 20);
    for (int w = 0; w < p->nstep; ++w)
      hipLaunchKernelGGL(plain_accum_kernel, dim3((unsigned)grid), dim3(256), 0, st,
                         d_spec + w * win, p->nspec, p->d_row_src + r0, p->d_row_rec + r0, nrows,
                         dst, p->xpitch, nbk, scale);
  } else if (p->precision == FXC_PREC_FP32) {
    hipLaunchKernelGGL(contract_fp32_kernel, dim3(nwg), dim3(256), 0, st, d_spec, p->nstep,
                       p->nsta, p->nspec, tiles, trows, dst, p->xpitch, (int)ntl, scale, c0, nbc);
  } else {
    const size_t lds = (size_t)kTilePairs * kBinChunk * sizeof(float2);
    // grid.y carries the tile count to the launcher (the launch itself is 1-D)
    FXC_TRY(launch_mfma_contract(p->precision, p->kh, dim3(nwg, ntl), lds, st, p->d_hi, p->d_lo,
                                 p->nsta, p->nspec, p->d_scale, tiles, trows, dst, p->xpitch,
                                 scale, c0, nbc));
  }
  FXC_HIP(hipGetLastError());
  FXC_TRY(xc_mark(p, kTimeContract, false, st));
  return FXC_OK;
}

// FXC_XC_LAG_R0=0 selects the production-length lag kernel without the
// fold in pass 0 (lag_kernel_fixed).
// large_fold_kernel of the one-window stack X[k] = conj(S[src][k]) S[rec][k]
// (rows r0.. of the product rows), formed from the spectra (rows of pitch) as
// they are read.
__global__ void large_fold_prod_kernel(const float2 *__restrict__ S, long long pitch,
                                       const int *__restrict__ src, const int *__restrict__ rec,
                                       int M, const float2 *__restrict__ half,
                                       float2 *__restrict__ Z) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const long long row = blockIdx.y;
  const float2 *ss = S + (size_t)src[row] * pitch, *sr = S + (size_t)rec[row] * pitch;
  auto x = [&](int i) {
    const float2 u = ss[i], v = sr[i];
    return make_float2(u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x);
  };
  float2 v;
  if (k == 0) {
    const float a = x(0).x, b = x(M).x;
    v = make_float2(a + b, a - b);
  } else {
    const float2 a = x(k);
    const float2 b = cconj(x(M - k));
    const float2 sv = cadd(a, b);
    const float2 d = cmul(half[k], csub(a, b));
    v = make_float2(sv.x - d.y, sv.y + d.x);
  }
  Z[row * M + k] = v;
}

static bool xc_lag_r0_on() {
  static const bool off = [] {
    const char *e = getenv("FXC_XC_LAG_R0");
    return e && atoi(e) == 0;
  }();
  return !off;
}

// task_scale (per task, may be null) multiplies the output scale of each row.
// prod_src/prod_rec (one window, the compile-time lag plan only): xs are the
// station spectra (rows of xpitch) and row r's stack is their product.
static int xc_lags_rows(fxc_xc_plan *p, const float2 *xs, long long xpitch, long long nrows,
                        const int *row_out, float *d_cc, hipStream_t st,
                        const float *task_scale = nullptr, const int *prod_src = nullptr,
                        const int *prod_rec = nullptr) {
  if (nrows == 0) return FXC_OK;
  FXC_CHECK_ARG(!prod_src || p->large_lag || (p->d_fix_tw && xc_lag_r0_on()),
                "xc: the one-window product form needs the compile-time or four-step lag");
  const double weight0 = 1.0 / (double)((float)p->nfft * p->dt);
  if (p->large_lag) {
    // four-step inverse of the folded spectrum, kLagLargeChunk rows at a time
    const int M = p->nfft / 2;
    const LargeFft *lp = nullptr;
    FXC_TRY(large_plan(M, 1, &lp));
    if (!p->d_lwork) FXC_TRY(dmalloc((void **)&p->d_lwork, (size_t)kLagLargeChunk * 2 * M * sizeof(float2)));
    float2 *w0 = p->d_lwork, *w1 = p->d_lwork + (size_t)kLagLargeChunk * M;
    FXC_TRY(xc_mark(p, kTimeLag, true, st));
    for (long long r0 = 0; r0 < nrows; r0 += kLagLargeChunk) {
      const long long nr = std::min<long long>(kLagLargeChunk, nrows - r0);
      if (prod_src)
        hipLaunchKernelGGL(large_fold_prod_kernel, dim3((unsigned)ceil_div(M, 256), (unsigned)nr),
                           dim3(256), 0, st, xs, xpitch, prod_src + r0, prod_rec + r0, M,
                           lp->d_half, w0);
      else
        hipLaunchKernelGGL(large_fold_kernel, dim3((unsigned)ceil_div(M, 256), (unsigned)nr),
                           dim3(256), 0, st, xs + r0 * xpitch, xpitch, M, lp->d_half, w0);
      FXC_HIP(hipGetLastError());
      FXC_TRY(large_c2c(lp, w0, M, w0, M, nr, w1, 1.f, st));
      hipLaunchKernelGGL(lag_extract_kernel,
                         dim3((unsigned)ceil_div(p->cc_size, 256), (unsigned)std::min(nr, 65535LL)),
                         dim3(256), 0, st, w0, M, nr, row_out ? row_out + r0 : nullptr,
                         row_out ? d_cc : d_cc + (size_t)r0 * p->cc_size, p->half_cc, weight0,
                         task_scale);
      FXC_HIP(hipGetLastError());
    }
    FXC_TRY(xc_mark(p, kTimeLag, false, st));
    return FXC_OK;
  }
  const RealFftPlan *rp = nullptr;
  FXC_TRY(plan_cache_real(p->nfft, 1, &rp, p->lag_max_radix, true));
  const size_t lds = (size_t)real_lds_count(p->nfft / 2) * sizeof(float2);
  auto kern = p->lag_max_radix >= 9 ? lag_kernel<kRadixLag9> : lag_kernel<kRadixLag8>;
  FXC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  int dev = 0, ncu = 0;
  FXC_HIP(hipGetDevice(&dev));
  FXC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const long long grid = std::min<long long>(nrows, (long long)std::max(ncu, 1));
  // reference weight: 1.0 / (width * dt) with width converted to float
  const double weight = 1.0 / (double)((float)p->nfft * p->dt);
  FXC_TRY(xc_mark(p, kTimeLag, true, st));
  if (p->d_fix_tw && xc_lag_r0_on()) {
    using FP = LagPlan18522;
    static_assert(lag_r0_lds_count<FP>() <= kLdsComplexCapacity, "lag LDS fits one CU");
    const size_t flds = (size_t)lag_r0_lds_count<FP>() * sizeof(float2);
    // the edge form of pass P-2 holds when both lag ends stay within 2G bins
    constexpr int G2 = 2 * fixed_edge_group<FP>();
    const int n_neg = (p->half_cc + 1) / 2, n_pos = p->half_cc / 2;
    const bool edge = n_pos < G2 && n_neg <= G2;
    auto fk = prod_src ? (edge ? lag_kernel_fixed_r0<FP, true, true>
                               : lag_kernel_fixed_r0<FP, false, true>)
                       : (edge ? lag_kernel_fixed_r0<FP, true> : lag_kernel_fixed_r0<FP, false>);
    FXC_HIP(hipFuncSetAttribute((const void *)fk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)flds));
    hipLaunchKernelGGL(fk, dim3((unsigned)grid), dim3(kLagThreads), flds, st, xs, xpitch, nrows,
                       row_out, d_cc, p->half_cc, weight, task_scale, p->d_fix_tw, p->d_fix_twl,
                       p->d_fix_fold, p->d_fix_roots, prod_src, prod_rec);
  } else if (p->d_fix_tw) {
    using FP = LagPlan18522;
    const size_t flds = (size_t)(FP::kM + 1 + fixed_lds_tw_count<FP>()) * sizeof(float2);
    auto fk = lag_kernel_fixed<FP>;
    FXC_HIP(hipFuncSetAttribute((const void *)fk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)flds));
    hipLaunchKernelGGL(fk, dim3((unsigned)grid), dim3(kLagThreads), flds, st, xs, xpitch, nrows,
                       row_out, d_cc, p->half_cc, weight, task_scale, p->d_fix_tw, p->d_fix_twl,
                       p->d_fix_fold, p->d_fix_roots);
  } else {
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kLagThreads), lds, st, xs, xpitch,
                       p->nspec, nrows, row_out, d_cc, p->half_cc, weight, task_scale, rp->half.dev,
                       rp->half.d_tw, rp->d_fold);
  }
  FXC_HIP(hipGetLastError());
  FXC_TRY(xc_mark(p, kTimeLag, false, st));
  return FXC_OK;
}

// Cross-spectra only (the windowed stack), rows in task order:
// d_xspec[task][k] = (1/nstep) sum_w conj(S[w][src][k]) S[w][rec][k].
extern "C" int fxc_xc_cross_spectra(fxc_xc_plan *p, const float *d_spec, float *d_xspec,
                                    void *stream) {
  FXC_CHECK_ARG(p && d_spec && d_xspec, "fxc_xc_cross_spectra: invalid arguments");
  FXC_CHECK_ARG(p->chunk_tile_begin.size() >= 1, "fxc_xc_cross_spectra: no pairs set");
  if (p->npairs == 0) return FXC_OK;
  hipStream_t st = (hipStream_t)stream;
  long long maxrows = 0;
  FXC_TRY(fxc_xc_pairs_info(p, nullptr, nullptr, &maxrows));
  FXC_TRY(xc_ensure_workspace(p, (size_t)maxrows));
  FXC_TRY(xc_band_prepare(p, st));
  FXC_TRY(xc_prepare_spectra(p, (const float2 *)d_spec, st));
  const long long nc = (long long)p->chunk_tile_begin.size() - 1;
  for (long long c = 0; c < nc; ++c) {
    FXC_TRY(xc_contract_chunk(p, (const float2 *)d_spec, c, p->d_xspec, st));
    const long long r0 = p->chunk_row_begin[c], r1 = p->chunk_row_begin[c + 1];
    if (r1 > r0) {
      hipLaunchKernelGGL(scatter_rows_kernel, dim3((unsigned)(r1 - r0)), dim3(256), 0, st,
                         p->d_xspec, p->xpitch, p->d_row_task + r0, p->nspec, (float2 *)d_xspec);
      FXC_HIP(hipGetLastError());
    }
  }
  return FXC_OK;
}

// Lag synthesis only: cc[r] = lags of the cross-spectrum row r.
extern "C" int fxc_xc_lags(fxc_xc_plan *p, const float *d_xspec, long long nrows, float *d_cc,
                           void *stream) {
  FXC_CHECK_ARG(p && d_xspec && d_cc && nrows >= 0, "fxc_xc_lags: invalid arguments");
  return xc_lags_rows(p, (const float2 *)d_xspec, p->nspec, nrows, nullptr, d_cc,
                      (hipStream_t)stream);
}

// Full correlation: cc[task][0..2H] for every task set by fxc_xc_set_pairs.
// One window at the compile-time lag length or a four-step one, every bin: a
// pair's stack is the product conj(S[src]) S[rec] of its two spectra, formed
// in FP32 as the lag synthesis reads them (the plan's operand form is not
// used: there is no sum over windows to contract).
static bool xc_one_window_product(const fxc_xc_plan *p) {
  return p->nstep == 1 && p->precision != FXC_PREC_PLAIN && p->band_c1 < 0 &&
         (p->large_lag || (p->d_fix_tw && xc_lag_r0_on())) && p->d_row_src && p->d_row_rec;
}

extern "C" int fxc_xc_correlate(fxc_xc_plan *p, const float *d_spec, float *d_cc, void *stream) {
  FXC_CHECK_ARG(p && d_spec && d_cc, "fxc_xc_correlate: invalid arguments");
  FXC_CHECK_ARG(p->chunk_tile_begin.size() >= 1, "fxc_xc_correlate: no pairs set");
  if (p->npairs == 0) return FXC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (xc_one_window_product(p))
    return xc_lags_rows(p, (const float2 *)d_spec, p->nspec, p->npairs, p->d_row_task, d_cc, st,
                        nullptr, p->d_row_src, p->d_row_rec);
  long long maxrows = 0;
  FXC_TRY(fxc_xc_pairs_info(p, nullptr, nullptr, &maxrows));
  FXC_TRY(xc_ensure_workspace(p, (size_t)maxrows));
  FXC_TRY(xc_band_prepare(p, st));
  FXC_TRY(xc_prepare_spectra(p, (const float2 *)d_spec, st));
  const long long nc = (long long)p->chunk_tile_begin.size() - 1;
  for (long long c = 0; c < nc; ++c) {
    FXC_TRY(xc_contract_chunk(p, (const float2 *)d_spec, c, p->d_xspec, st));
    const long long r0 = p->chunk_row_begin[c], r1 = p->chunk_row_begin[c + 1];
    FXC_TRY(xc_lags_rows(p, p->d_xspec, p->xpitch, r1 - r0, p->d_row_task + r0, d_cc, st));
  }
  return FXC_OK;
}

extern "C" int fxc_xc_set_band(fxc_xc_plan *p, int k_lo, int k_hi) {
  FXC_CHECK_ARG(p, "fxc_xc_set_band: null plan");
  FXC_CHECK_ARG(k_lo >= 0 && k_lo <= k_hi && k_hi < p->nspec,
                "fxc_xc_set_band: band [%d, %d] outside bins [0, %d)", k_lo, k_hi, p->nspec);
  if (k_lo == 0 && k_hi == p->nspec - 1) {
    p->band_c0 = 0;
    p->band_c1 = -1;
    return FXC_OK;
  }
  if (p->precision == FXC_PREC_PLAIN) {
    set_error("fxc_xc_set_band: the PLAIN form contracts every bin (parity only)");
    return FXC_ERR_UNSUPPORTED;
  }
  p->band_c0 = k_lo / kBinChunk;
  p->band_c1 = ceil_div(k_hi + 1, kBinChunk);
  p->band_zero_rows = 0;
  return FXC_OK;
}

extern "C" int fxc_xc_plan_info(const fxc_xc_plan *p, int *nfft, int *half_cc, int *cc_size) {
  FXC_CHECK_ARG(p, "fxc_xc_plan_info: null plan");
  if (nfft) *nfft = p->nfft;
  if (half_cc) *half_cc = p->half_cc;
  if (cc_size) *cc_size = p->cc_size;
  return FXC_OK;
}

// ---------------------------------------------------------------------------
// Uneven time segments in one set of launches.
//
// Segment s holds nsta[s] stations and nstep[s] windows; its pairs are all
// (i, j), i < j, of its stations, in row-major order, and the batch's output
// rows are the segments' pairs one segment after another.  Segments share
// plans ("groups"): the stations of a group's segments sit side by side on
// one operand axis, each segment starting on a 32-station tile boundary, so
// every contraction tile belongs to one segment and a group runs one split
// launch, one persistent contraction and one lag launch per row chunk.  A
// group's operand depth K is its plan's kh, fixed by its largest window
// count; segments with fewer windows carry zeros in the rest.  Grouping 0
// puts every segment in one group (fewest launches, deepest K); grouping 1
// makes one group per compiled operand depth (the window count rounded up
// to 8, 16, 24, 32, 40, 48, 64, 96 or 128, as xc_operand_depth does), so
// short segments are not padded to the longest.  fxc_xc_batch_tune times
// both on the caller's data and keeps the faster.
//
// The contraction of a group divides by the group's window count; each row's
// output scale then carries group_nstep / divisor, where the divisor is the
// segment's window count (gap mode KEEP, the reference's rule: zero-filled
// station-windows still count) or, per pair, the number of windows in which
// both stations are valid (gap mode EXCLUDE).  Station-windows whose valid
// byte is 0 enter the contraction as zeros in both modes.
// ---------------------------------------------------------------------------
__global__ void batch_task_scale_kernel(const int4 *__restrict__ tinfo,
                                        const SplitSeg *__restrict__ segs,
                                        const unsigned char *__restrict__ valid, long long ntasks,
                                        float *__restrict__ task_scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntasks) return;
  const int4 ti = tinfo[t];  // segment, local src, local rec, group window count
  const SplitSeg sg = segs[ti.x];
  const unsigned char *v = valid + sg.valid_off;
  int cnt = 0;
  for (int w = 0; w < sg.nstep; ++w)
    cnt += (v[(size_t)w * sg.nsta + ti.y] != 0 && v[(size_t)w * sg.nsta + ti.z] != 0) ? 1 : 0;
  task_scale[t] = cnt > 0 ? (float)((double)ti.w / (double)cnt) : 0.f;
}

struct XcBatchGroup {
  fxc_xc_plan *plan = nullptr;
  std::vector<int> segs;
  int max_nsta = 0;
  int nsta_dst = 0;
  SplitSeg *d_segs = nullptr;  // the group's segments, sta_off on its own axis
  int *d_row_out = nullptr;    // plan workspace row -> batch task
};

struct fxc_xc_batch {
  int nspec = 0, half_cc = 0, precision = FXC_PREC_BF16X3, gap_mode = FXC_GAPS_KEEP;
  float dt = 0.f;
  long long max_rows = 0;
  int grouping = 0;
  std::vector<int> nsta, nstep;
  std::vector<long long> task_base;  // first batch task of each segment
  long long ntasks = 0, spec_elems = 0, valid_elems = 0;
  std::vector<SplitSeg> segs;  // batch offsets; sta_off unused here
  std::vector<XcBatchGroup> groups;
  SplitSeg *d_segs = nullptr;
  int4 *d_tinfo = nullptr;
  float *d_scale_keep = nullptr;  // KEEP-mode row scales (fixed by the shapes)
  float *d_scale = nullptr;       // EXCLUDE-mode row scales (from the valid bytes)
  // one cross-spectrum workspace and one pair of split operands for all the
  // groups: their plans run one after another on one stream
  float2 *d_ws_x = nullptr;
  unsigned short *d_ws_hi = nullptr, *d_ws_lo = nullptr;
  size_t ws_rows = 0, ws_elems = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int batch_depth(int nstep) { return xc_operand_depth(nstep); }

static void batch_free_groups(fxc_xc_batch *b) {
  for (auto &g : b->groups) {
    fxc_xc_plan_destroy(g.plan);
    dfree(g.d_segs);
    dfree(g.d_row_out);
  }
  b->groups.clear();
  dfree(b->d_tinfo);
  dfree(b->d_scale_keep);
  b->d_tinfo = nullptr;
  b->d_scale_keep = nullptr;
}

static void batch_drop_graph(fxc_xc_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// (Re)builds the groups of the current grouping: plans, tiles, row maps and
// the per-task tables.
static int batch_build(fxc_xc_batch *b) {
  batch_drop_graph(b);
  batch_free_groups(b);
  const int nseg = (int)b->nsta.size();
  std::map<int, int> by_kh;  // K depth -> group index (grouping 1)
  std::vector<int> seg_group(nseg);
  for (int s = 0; s < nseg; ++s) {
    const int key = b->grouping ? batch_depth(b->nstep[s]) : 0;
    auto it = by_kh.find(key);
    if (it == by_kh.end()) {
      it = by_kh.emplace(key, (int)b->groups.size()).first;
      b->groups.emplace_back();
    }
    seg_group[s] = it->second;
    b->groups[it->second].segs.push_back(s);
  }
  std::vector<int4> tinfo((size_t)b->ntasks);
  std::vector<float> keep((size_t)b->ntasks);
  for (auto &g : b->groups) {
    int gstep = 1;
    for (int s : g.segs) gstep = std::max(gstep, b->nstep[s]);
    std::vector<SplitSeg> gs;
    std::vector<int> src, rec;
    std::vector<long long> task;  // plan task -> batch task
    int off = 0;
    for (int s : g.segs) {
      SplitSeg d = b->segs[s];
      d.sta_off = off;
      gs.push_back(d);
      long long t = b->task_base[s];
      for (int i = 0; i < b->nsta[s]; ++i)
        for (int j = i + 1; j < b->nsta[s]; ++j, ++t) {
          src.push_back(off + i);
          rec.push_back(off + j);
          task.push_back(t);
          tinfo[(size_t)t] = make_int4(s, i, j, gstep);
          keep[(size_t)t] = (float)((double)gstep / (double)b->nstep[s]);
        }
      g.max_nsta = std::max(g.max_nsta, b->nsta[s]);
      off += (b->nsta[s] + kTile - 1) / kTile * kTile;
    }
    g.nsta_dst = std::max(off, 1);
    FXC_TRY(fxc_xc_plan_create(&g.plan, b->nspec, gstep, b->dt, b->half_cc, b->precision));
    FXC_TRY(fxc_xc_set_chunk_rows(g.plan, b->max_rows));
    FXC_TRY(fxc_xc_set_pairs(g.plan, g.nsta_dst, src.data(), rec.data(), (long long)src.size()));
    std::vector<int> row_out(g.plan->h_row_task.size());
    for (size_t r = 0; r < row_out.size(); ++r)
      row_out[r] = (int)task[(size_t)g.plan->h_row_task[r]];
    FXC_TRY(dmalloc((void **)&g.d_segs, gs.size() * sizeof(SplitSeg)));
    FXC_HIP(hipMemcpy(g.d_segs, gs.data(), gs.size() * sizeof(SplitSeg), hipMemcpyHostToDevice));
    if (!row_out.empty()) {
      FXC_TRY(dmalloc((void **)&g.d_row_out, row_out.size() * sizeof(int)));
      FXC_HIP(hipMemcpy(g.d_row_out, row_out.data(), row_out.size() * sizeof(int),
                        hipMemcpyHostToDevice));
    }
  }
  if (b->ntasks > 0) {
    FXC_TRY(dmalloc((void **)&b->d_tinfo, tinfo.size() * sizeof(int4)));
    FXC_HIP(hipMemcpy(b->d_tinfo, tinfo.data(), tinfo.size() * sizeof(int4),
                      hipMemcpyHostToDevice));
    FXC_TRY(dmalloc((void **)&b->d_scale_keep, keep.size() * sizeof(float)));
    FXC_HIP(hipMemcpy(b->d_scale_keep, keep.data(), keep.size() * sizeof(float),
                      hipMemcpyHostToDevice));
  }
  return FXC_OK;
}

extern "C" int fxc_xc_batch_create(fxc_xc_batch **out, int nspec, float dt, int half_cc,
                                   int precision, int nseg, const int *nsta, const int *nstep,
                                   int gap_mode, long long max_rows) {
  FXC_CHECK_ARG(out != nullptr, "fxc_xc_batch_create: null output");
  *out = nullptr;
  FXC_CHECK_ARG(nseg >= 1 && nseg <= 65535 && nsta && nstep,
                "fxc_xc_batch_create: need 1..65535 segments");
  FXC_CHECK_ARG(gap_mode == FXC_GAPS_KEEP || gap_mode == FXC_GAPS_EXCLUDE,
                "fxc_xc_batch_create: unknown gap mode %d", gap_mode);
  FXC_CHECK_ARG(max_rows >= 0, "fxc_xc_batch_create: negative row cap");
  if (precision != FXC_PREC_BF16X3) {
    set_error("fxc_xc_batch_create: the batched form runs the split-BF16 contraction only");
    return FXC_ERR_UNSUPPORTED;
  }
  auto *b = new (std::nothrow) fxc_xc_batch();
  if (!b) {
    set_error("fxc_xc_batch_create: out of host memory");
    return FXC_ERR_ALLOC;
  }
  b->nspec = nspec;
  b->dt = dt;
  b->half_cc = half_cc;
  b->precision = precision;
  b->gap_mode = gap_mode;
  b->max_rows = max_rows;
  long long soff = 0, voff = 0, t = 0;
  for (int s = 0; s < nseg; ++s) {
    if (nsta[s] < 1 || nstep[s] < 1 || nstep[s] > 128) {
      set_error("fxc_xc_batch_create: segment %d has %d stations and %d windows (need >= 1 and "
                "<= 128 windows)", s, nsta[s], nstep[s]);
      delete b;
      return FXC_ERR_ARG;
    }
    b->nsta.push_back(nsta[s]);
    b->nstep.push_back(nstep[s]);
    b->task_base.push_back(t);
    SplitSeg d{};
    d.spec_off = soff;
    d.valid_off = voff;
    d.nsta = nsta[s];
    d.nstep = nstep[s];
    d.npairs = (int)((long long)nsta[s] * (nsta[s] - 1) / 2);
    b->segs.push_back(d);
    soff += (long long)nstep[s] * nsta[s] * nspec;
    voff += (long long)nstep[s] * nsta[s];
    t += (long long)nsta[s] * (nsta[s] - 1) / 2;
  }
  b->spec_elems = soff;
  b->valid_elems = voff;
  b->ntasks = t;
  if (t > INT_MAX) {
    set_error("fxc_xc_batch_create: %lld pairs exceed the row index range", t);
    delete b;
    return FXC_ERR_ARG;
  }
  int rc = dmalloc((void **)&b->d_segs, b->segs.size() * sizeof(SplitSeg));
  if (rc == FXC_OK && hipMemcpy(b->d_segs, b->segs.data(), b->segs.size() * sizeof(SplitSeg),
                                hipMemcpyHostToDevice) != hipSuccess) {
    set_error("fxc_xc_batch_create: segment table upload failed");
    rc = FXC_ERR_HIP;
  }
  if (rc == FXC_OK && b->ntasks > 0)
    rc = dmalloc((void **)&b->d_scale, (size_t)b->ntasks * sizeof(float));
  if (rc == FXC_OK) rc = batch_build(b);
  if (rc != FXC_OK) {
    fxc_xc_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FXC_OK;
}

extern "C" int fxc_xc_batch_destroy(fxc_xc_batch *b) {
  if (!b) return FXC_OK;
  batch_drop_graph(b);
  batch_free_groups(b);
  dfree(b->d_segs);
  dfree(b->d_scale);
  dfree(b->d_ws_x);
  dfree(b->d_ws_hi);
  dfree(b->d_ws_lo);
  delete b;
  return FXC_OK;
}

extern "C" int fxc_xc_batch_info(const fxc_xc_batch *b, long long *ntasks, long long *spec_elems,
                                 long long *valid_elems, int *ngroups, int *grouping) {
  FXC_CHECK_ARG(b, "fxc_xc_batch_info: null batch");
  if (ntasks) *ntasks = b->ntasks;
  if (spec_elems) *spec_elems = b->spec_elems;
  if (valid_elems) *valid_elems = b->valid_elems;
  if (ngroups) *ngroups = (int)b->groups.size();
  if (grouping) *grouping = b->grouping;
  return FXC_OK;
}

extern "C" int fxc_xc_batch_set_grouping(fxc_xc_batch *b, int grouping) {
  FXC_CHECK_ARG(b && (grouping == 0 || grouping == 1),
                "fxc_xc_batch_set_grouping: grouping must be 0 or 1");
  if (grouping == b->grouping && !b->groups.empty()) return FXC_OK;
  b->grouping = grouping;
  return batch_build(b);
}

// Row cap of each plan's cross-spectrum workspace (0: no cap); rebuilds the
// groups, so a captured graph is dropped.
extern "C" int fxc_xc_batch_set_rows(fxc_xc_batch *b, long long max_rows) {
  FXC_CHECK_ARG(b && max_rows >= 0, "fxc_xc_batch_set_rows: invalid arguments");
  if (max_rows == b->max_rows && !b->groups.empty()) return FXC_OK;
  b->max_rows = max_rows;
  return batch_build(b);
}

// The shared workspaces, sized for the largest group (grow-only), lent to
// every group's plan.
static int batch_share_workspace(fxc_xc_batch *b) {
  size_t rows = 0, elems = 0, pitch = 0;
  for (auto &g : b->groups) {
    fxc_xc_plan *p = g.plan;
    pitch = std::max(pitch, (size_t)p->xpitch);
    if (p->npairs == 0) continue;
    long long maxrows = 0;
    FXC_TRY(fxc_xc_pairs_info(p, nullptr, nullptr, &maxrows));
    rows = std::max(rows, (size_t)maxrows);
    elems = std::max(elems, (size_t)p->nsta * p->nspec * (size_t)(2 * p->kh));
  }
  if (rows > b->ws_rows) {
    dfree(b->d_ws_x);
    b->d_ws_x = nullptr;
    b->ws_rows = 0;
    FXC_TRY(dmalloc((void **)&b->d_ws_x, rows * pitch * sizeof(float2)));
    b->ws_rows = rows;
  }
  if (elems > b->ws_elems) {
    dfree(b->d_ws_hi);
    dfree(b->d_ws_lo);
    b->d_ws_hi = b->d_ws_lo = nullptr;
    b->ws_elems = 0;
    FXC_TRY(dmalloc((void **)&b->d_ws_hi, elems * sizeof(unsigned short)));
    FXC_TRY(dmalloc((void **)&b->d_ws_lo, elems * sizeof(unsigned short)));
    b->ws_elems = elems;
  }
  for (auto &g : b->groups) {
    fxc_xc_plan *p = g.plan;
    if (!p->ws_shared) {
      dfree(p->d_xspec);
      dfree(p->d_hi);
      dfree(p->d_lo);
    }
    p->d_xspec = b->d_ws_x;
    p->xspec_rows = b->ws_rows;
    p->d_hi = b->d_ws_hi;
    p->d_lo = b->d_ws_lo;
    p->split_elems = b->ws_elems;
    p->ws_shared = true;
  }
  return FXC_OK;
}

// Every launch of one batch run on `st`; allocates nothing once the shared
// workspaces exist (the first run creates them).
static int batch_run(fxc_xc_batch *b, const float2 *spec, const unsigned char *valid, float *cc,
                     hipStream_t st) {
  FXC_TRY(batch_share_workspace(b));
  const float *scale = b->d_scale_keep;
  if (b->gap_mode == FXC_GAPS_EXCLUDE && valid && b->ntasks > 0) {
    hipLaunchKernelGGL(batch_task_scale_kernel, dim3((unsigned)ceil_div(b->ntasks, 256)),
                       dim3(256), 0, st, b->d_tinfo, b->d_segs, valid, b->ntasks, b->d_scale);
    FXC_HIP(hipGetLastError());
    scale = b->d_scale;
  }
  for (auto &g : b->groups) {
    fxc_xc_plan *p = g.plan;
    if (p->npairs == 0) continue;
    long long maxrows = 0;
    FXC_TRY(fxc_xc_pairs_info(p, nullptr, nullptr, &maxrows));
    FXC_TRY(xc_ensure_workspace(p, (size_t)maxrows));
    const dim3 grid((unsigned)ceil_div(p->nspec, kSplitBins),
                    (unsigned)ceil_div(g.max_nsta, kSplitSta), (unsigned)g.segs.size());
    hipLaunchKernelGGL(split_segments_kernel<FXC_PREC_BF16X3>, grid, dim3(kSplitThreads), 0, st, spec,
                       valid, g.d_segs, g.nsta_dst, p->nspec, p->kh, p->d_hi, p->d_lo);
    FXC_HIP(hipGetLastError());
    const long long nc = (long long)p->chunk_tile_begin.size() - 1;
    for (long long c = 0; c < nc; ++c) {
      FXC_TRY(xc_contract_chunk(p, spec, c, p->d_xspec, st));
      const long long r0 = p->chunk_row_begin[c], r1 = p->chunk_row_begin[c + 1];
      FXC_TRY(xc_lags_rows(p, p->d_xspec, p->xpitch, r1 - r0, g.d_row_out + r0, cc, st, scale));
    }
  }
  return FXC_OK;
}

static int batch_check_io(const fxc_xc_batch *b, const float *d_spec, const float *d_cc,
                          const char *who) {
  FXC_CHECK_ARG(b && d_spec && d_cc, "%s: invalid arguments", who);
  return FXC_OK;
}

// cc[task][0..2H] for every pair of every segment.  d_spec: the segments'
// spectra [nstep][nsta][nspec] complex, one after another; d_valid: null
// (every station-window valid) or bytes [nstep][nsta], one segment after
// another.
extern "C" int fxc_xc_batch_correlate(fxc_xc_batch *b, const float *d_spec,
                                      const unsigned char *d_valid, float *d_cc, void *stream) {
  FXC_TRY(batch_check_io(b, d_spec, d_cc, "fxc_xc_batch_correlate"));
  return batch_run(b, (const float2 *)d_spec, d_valid, d_cc, (hipStream_t)stream);
}

// Records one batch run on `stream` (not the null stream) into a HIP graph
// bound to these buffers, after one eager run that creates every workspace.
// fxc_xc_batch_replay launches it again; new data are written into the same
// buffers between replays.
extern "C" int fxc_xc_batch_capture(fxc_xc_batch *b, const float *d_spec,
                                    const unsigned char *d_valid, float *d_cc, void *stream) {
  FXC_TRY(batch_check_io(b, d_spec, d_cc, "fxc_xc_batch_capture"));
  FXC_CHECK_ARG(stream != nullptr, "fxc_xc_batch_capture: capture needs a non-null stream");
  hipStream_t st = (hipStream_t)stream;
  batch_drop_graph(b);
  FXC_TRY(batch_run(b, (const float2 *)d_spec, d_valid, d_cc, st));
  FXC_HIP(hipStreamSynchronize(st));
  FXC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  const int rc = batch_run(b, (const float2 *)d_spec, d_valid, d_cc, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != FXC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  FXC_HIP(e);
  b->graph = g;
  FXC_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return FXC_OK;
}

extern "C" int fxc_xc_batch_replay(fxc_xc_batch *b, void *stream) {
  FXC_CHECK_ARG(b && b->exec, "fxc_xc_batch_replay: no captured graph");
  FXC_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return FXC_OK;
}

// Runs the batch under both groupings (one warm-up run and `reps` timed runs
// each, HIP events on `stream`) and keeps the faster; ms[2] receives the
// mean time per run of grouping 0 and 1.  Any captured graph is dropped.
extern "C" int fxc_xc_batch_tune(fxc_xc_batch *b, const float *d_spec,
                                 const unsigned char *d_valid, float *d_cc, int reps,
                                 void *stream, double *ms) {
  FXC_TRY(batch_check_io(b, d_spec, d_cc, "fxc_xc_batch_tune"));
  FXC_CHECK_ARG(reps >= 1, "fxc_xc_batch_tune: reps must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  double t[2] = {0.0, 0.0};
  for (int mode = 0; mode < 2; ++mode) {
    FXC_TRY(fxc_xc_batch_set_grouping(b, mode));
    FXC_TRY(batch_run(b, (const float2 *)d_spec, d_valid, d_cc, st));
    hipEvent_t e0, e1;
    FXC_HIP(hipEventCreate(&e0));
    FXC_HIP(hipEventCreate(&e1));
    int rc = hipEventRecord(e0, st) == hipSuccess ? FXC_OK : FXC_ERR_HIP;
    for (int r = 0; r < reps && rc == FXC_OK; ++r)
      rc = batch_run(b, (const float2 *)d_spec, d_valid, d_cc, st);
    float el = 0.f;
    if (rc == FXC_OK && (hipEventRecord(e1, st) != hipSuccess ||
                         hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&el, e0, e1) != hipSuccess)) {
      set_error("fxc_xc_batch_tune: event timing failed");
      rc = FXC_ERR_HIP;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    FXC_TRY(rc);
    t[mode] = (double)el / reps;
  }
  if (ms) {
    ms[0] = t[0];
    ms[1] = t[1];
  }
  return fxc_xc_batch_set_grouping(b, t[1] < t[0] ? 1 : 0);
}
