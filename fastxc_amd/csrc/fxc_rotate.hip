// ENZ -> RTZ rotation of the nine component correlations of a station pair.
//
// Replaces rotate_enz_to_rtz / build_rotate_matrix
// (fastxc/operators/rotation/python_rotate.py:34-104), which the reference
// applies pair by pair in numpy after stacking (:264-390).  The source-side
// horizontal basis is (R, T) = (E sin az + N cos az, E cos az - N sin az), the
// receiver-side one (R, T) = (-E sin baz - N cos baz, -E cos baz + N sin baz)
// (radial pointing from the receiver back to the source), Z passes through;
// the tensor component XY of the pair is sum_ij s_X[i] r_Y[j] C_ij over the
// horizontal components i, j in {E, N}.  Each matrix entry is one float
// product of the caller's float sines and cosines, as in the reference.
//
// MI355X design: an elementwise, HBM-bound map (36 bytes read and 36 written
// per sample and pair).  A workgroup covers 256 consecutive samples of one
// pair, so each of the nine component rows is read and written as one
// coalesced 1 KiB run; every thread holds its samples' nine inputs in
// registers before it writes, so the rotation can run in place.  Two samples
// per thread keep 18 loads in flight: 4.66 TB/s against 4.85 TB/s for the
// runtime's device-to-device copy of the same bytes.
#include "fxc_common.hpp"
#include "fastxc_amd.h"

namespace fxc {

constexpr int kRotThreads = 256;
constexpr int kRotPer = 2;  // samples per thread, kRotThreads apart

// enz/rtz [npairs][9][n] in the order EE EN EZ NE NN NZ ZE ZN ZZ /
// RR RT RZ TR TT TZ ZR ZT ZZ; trig [npairs] (sin az, cos az, sin baz, cos baz).
// A workgroup covers kRotPer * kRotThreads samples of one pair; all loads of
// a thread are issued before its first store (18 in flight).  enz and rtz
// may be one buffer: each sample is read and then written by one thread.
__global__ __launch_bounds__(kRotThreads) void rotate_rtz_kernel(
    const float *enz, const float4 *__restrict__ trig, int n, int chunks, float *rtz) {
  const long long p = blockIdx.x / chunks;
  const int t0 = (int)(blockIdx.x % chunks) * kRotThreads * kRotPer + (int)threadIdx.x;
  const float4 g = trig[p];
  // horizontal bases over (E, N): s_R s_T at the source, r_R r_T at the receiver
  const float s[2][2] = {{g.x, g.y}, {g.y, -g.x}};
  const float r[2][2] = {{-g.z, -g.w}, {-g.w, g.z}};
  const size_t base = (size_t)p * 9 * (size_t)n;
  float c[kRotPer][9];
#pragma unroll
  for (int q = 0; q < kRotPer; ++q) {
    const int t = t0 + q * kRotThreads;
    if (t < n) {
#pragma unroll
      for (int k = 0; k < 9; ++k) c[q][k] = enz[base + (size_t)k * n + t];
    }
  }
#pragma unroll
  for (int q = 0; q < kRotPer; ++q) {
    const int t = t0 + q * kRotThreads;
    if (t >= n) continue;
    float y[9];
    // horizontal-horizontal: XY = sum over i, j of s_X[i] r_Y[j] C_(i,j),
    // in the order EE, EN, NE, NN (C at 0, 1, 3, 4)
#pragma unroll
    for (int X = 0; X < 2; ++X)
#pragma unroll
      for (int Y = 0; Y < 2; ++Y) {
        float acc = (s[X][0] * r[Y][0]) * c[q][0];
        acc = fmaf(s[X][0] * r[Y][1], c[q][1], acc);
        acc = fmaf(s[X][1] * r[Y][0], c[q][3], acc);
        acc = fmaf(s[X][1] * r[Y][1], c[q][4], acc);
        y[X * 3 + Y] = acc;
      }
    // horizontal-Z: XZ = s_X[E] EZ + s_X[N] NZ; Z-horizontal: ZY = r_Y[E] ZE + r_Y[N] ZN
#pragma unroll
    for (int X = 0; X < 2; ++X) y[X * 3 + 2] = fmaf(s[X][1], c[q][5], s[X][0] * c[q][2]);
#pragma unroll
    for (int Y = 0; Y < 2; ++Y) y[6 + Y] = fmaf(r[Y][1], c[q][7], r[Y][0] * c[q][6]);
    y[8] = c[q][8];
#pragma unroll
    for (int k = 0; k < 9; ++k) rtz[base + (size_t)k * n + t] = y[k];
  }
}

}  // namespace fxc

using namespace fxc;

extern "C" int fxc_rotate_rtz(const float *d_enz, const float *d_trig, long long npairs, int n,
                              float *d_rtz, void *stream) {
  FXC_CHECK_ARG(d_enz && d_trig && d_rtz && npairs >= 1 && n >= 1,
                "fxc_rotate_rtz: invalid arguments");
  FXC_CHECK_ARG(((uintptr_t)d_trig & 15) == 0, "fxc_rotate_rtz: trig must be 16-byte aligned");
  const long long span = (long long)kRotThreads * kRotPer;
  const long long chunks = (n + span - 1) / span;
  FXC_CHECK_ARG(npairs * chunks <= 0x7fffffffLL,
                "fxc_rotate_rtz: %lld pairs of %d samples exceed one launch", npairs, n);
  hipLaunchKernelGGL(rotate_rtz_kernel, dim3((unsigned)(npairs * chunks)), dim3(kRotThreads), 0,
                     (hipStream_t)stream, d_enz, (const float4 *)d_trig, n, (int)chunks, d_rtz);
  FXC_HIP(hipGetLastError());
  return FXC_OK;
}
