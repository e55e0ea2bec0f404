/*
 * Multi-threaded C restatement of xc_fast's per-pair arithmetic -- TEST AND
 * BASELINE INFRASTRUCTURE ONLY (the CPU baseline of bench.py and a checker).
 *
 * For every pair (src, rec) of one timestamp:
 *   X[k] = (1/nstep) sum_w conj(S[w][src][k]) S[w][rec][k]        FP32, window order
 *        (accumulateStepXc2DKernel, native/xc/kernels/cuda.kernels.cu:16-36,
 *         launched per window at native/xc/worker/async_workers.cu:65-83)
 *   y    = C2R_nfft((-1)^k X) * 1/(nfft*dt), lags [nfft/2-H, nfft/2+H]
 *        (finalize_xc_batch, native/xc/worker/executor.cu:6-45)
 * The C2R runs as one complex FFT of length nfft/2 (half-length fold) with the
 * mixed-radix Stockham FFT of cpu_fft.h; lag m reads sample
 * (m - H) mod nfft of the unshifted C2R, which equals the reference's
 * (-1)^k shift.  Pairs are split over pthreads in contiguous blocks.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cpu_fft.h"

typedef struct {
  const cf *spec;
  int nstep, nsta, nspec, half_cc;
  float dt;
  const int *src, *rec;
  long begin, end;
  float *out;
  const fftplan *plan;
  const cf *fold; /* exp(+2 pi i k / nfft), k < M */
} job_t;

static void *worker(void *arg) {
  job_t *jb = (job_t *)arg;
  const int M = jb->nspec - 1, nfft = 2 * M, H = jb->half_cc, cc = 2 * H + 1;
  const double weight = 1.0 / (double)((float)nfft * jb->dt);
  const float scale = 1.0f / (float)jb->nstep;
  cf *X = (cf *)malloc(sizeof(cf) * (size_t)jb->nspec);
  cf *Z = (cf *)malloc(sizeof(cf) * (size_t)M);
  cf *W = (cf *)malloc(sizeof(cf) * (size_t)M);
  long pidx;
  for (pidx = jb->begin; pidx < jb->end; ++pidx) {
    const int s = jb->src[pidx], r = jb->rec[pidx];
    int k, w, m;
    memset(X, 0, sizeof(cf) * (size_t)jb->nspec);
    for (w = 0; w < jb->nstep; ++w) {
      const cf *a = jb->spec + ((size_t)w * jb->nsta + s) * jb->nspec;
      const cf *b = jb->spec + ((size_t)w * jb->nsta + r) * jb->nspec;
      for (k = 0; k < jb->nspec; ++k) {
        /* stack += conj(src) * rec * (1/nstep) */
        const float pr = a[k].re * b[k].re + a[k].im * b[k].im;
        const float pi = a[k].re * b[k].im - a[k].im * b[k].re;
        X[k].re += pr * scale;
        X[k].im += pi * scale;
      }
    }
    for (k = 0; k < M; ++k) {
      cf a = X[k], b = X[M - k];
      cf s1, d, wd;
      if (k == 0) {
        a.im = 0.f;
        b.im = 0.f;
      }
      b.im = -b.im;
      s1.re = a.re + b.re; s1.im = a.im + b.im;
      d.re = a.re - b.re; d.im = a.im - b.im;
      wd.re = jb->fold[k].re * d.re - jb->fold[k].im * d.im;
      wd.im = jb->fold[k].re * d.im + jb->fold[k].im * d.re;
      Z[k].re = s1.re - wd.im;
      Z[k].im = s1.im + wd.re;
    }
    fft_exec(jb->plan, Z, W);
    for (m = 0; m < cc; ++m) {
      int l = m - H, par = l & 1, n = (l - par) / 2;
      float v;
      if (n < 0) n += M;
      v = par ? Z[n].im : Z[n].re;
      jb->out[(size_t)pidx * cc + m] = (float)((double)v * weight);
    }
  }
  free(X);
  free(Z);
  free(W);
  return NULL;
}

/* spec: [nstep][nsta][nspec] complex64; out: [npairs][2*half_cc+1]. */
int cpu_xc_correlate(const float *spec, int nstep, int nsta, int nspec, const int *src,
                     const int *rec, long npairs, int half_cc, float dt, float *out,
                     int nthreads) {
  fftplan plan;
  const int M = nspec - 1;
  cf *fold;
  pthread_t th[256];
  job_t jobs[256];
  int t, k;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (fft_plan_init(&plan, M, +1) != 0) return -1;
  fold = (cf *)malloc(sizeof(cf) * (size_t)M);
  for (k = 0; k < M; ++k) {
    double a = 2.0 * M_PI * (double)k / (2.0 * M);
    fold[k].re = (float)cos(a);
    fold[k].im = (float)sin(a);
  }
  for (t = 0; t < nthreads; ++t) {
    jobs[t].spec = (const cf *)spec;
    jobs[t].nstep = nstep;
    jobs[t].nsta = nsta;
    jobs[t].nspec = nspec;
    jobs[t].half_cc = half_cc;
    jobs[t].dt = dt;
    jobs[t].src = src;
    jobs[t].rec = rec;
    jobs[t].begin = npairs * t / nthreads;
    jobs[t].end = npairs * (t + 1) / nthreads;
    jobs[t].out = out;
    jobs[t].plan = &plan;
    jobs[t].fold = fold;
    pthread_create(&th[t], NULL, worker, &jobs[t]);
  }
  for (t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(fold);
  fft_plan_free(&plan);
  return 0;
}
