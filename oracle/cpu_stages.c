/*
 * Multi-threaded C restatements of the stage arithmetic -- TEST AND BASELINE
 * INFRASTRUCTURE ONLY (the CPU baselines of bench.py's stage lines, checked
 * against the FP64 oracle by tests/test_cpu_stages.py).  FP32 data with the
 * reference's FP64 sums, the FFT of cpu_fft.h, work split over pthreads.
 *
 *   cpu_s2s_process   ComputeWorkerBatch (native/sac2spec/worker/compute.cu:8-272):
 *                     window cut, preprocess (core/preprocess.cu:23-105,
 *                     kernels/rdcrtr.cu:94-119), prefilter (compute.cu:59-86),
 *                     RUN-ABS or one-bit (core/normalization.cu:77-213),
 *                     whitening (core/spectrum.cu:178-305), output spectrum
 *                     (compute.cu:226-264); RUN-ABS-MF is not restated here
 *   cpu_pws           compute_pws_host_batch (native/pws/worker/gpu.cu:87-282,
 *                     kernels/cuda.pws.cu:3-95), power 1
 *   cpu_tfpws         compute_tfpws_from_prestack (native/tfpws/core/tfpws_compute.cu:19-366,
 *                     kernels/cuda.stransform.cu:59-332), every voice, power 1,
 *                     no band blend
 *   cpu_linear_stack  _linear_stack_group (fastxc/operators/stacking/linear_stack.py:228-247)
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "cpu_fft.h"

#define MINVAL 1e-30f
#define MAXVAL 20.0f

static int finite_cf(cf v) { return isfinite(v.re) && isfinite(v.im); }

typedef struct {
  int npts, seg, shift, nfft_f, nfft_o, nch, whiten, normalize, taper, runabs_win, whiten_win;
  int idx[4];
  float dt;
  const float *traces; /* [nfiles][npts] */
  int nfiles;
  const int *steps;
  int nsteps;
  const float *resp0; /* [nfft_f/2+1] */
  float *out;         /* [nsteps][nfiles][nfft_o/2+1] complex */
  const fftplan *pf_f, *pf_i, *ps_f, *ps_i, *po_f;
  cf *fold_f, *ifold_f, *fold_s, *ifold_s, *fold_o;
  long frame_begin, frame_end;
} s2s_ctx;

/* w[t] = sum_ch (1/win) sum_{i = t - win/2}^{t - win/2 + win - 1} v_ch[i] (edges
 * clipped), FP64 prefix sums per channel, channels added in float; clamped
 * to >= MINVAL (non-finite -> MINVAL). */
static void smooth_weights(const float *v, int nch, long pitch, int width, int win, float *w,
                           double *pre) {
  int ch, t;
  for (t = 0; t < width; ++t) w[t] = 0.f;
  for (ch = 0; ch < nch; ++ch) {
    const float *x = v + ch * pitch;
    pre[0] = 0.0;
    for (t = 0; t < width; ++t) pre[t + 1] = pre[t] + (double)x[t];
    for (t = 0; t < width; ++t) {
      int a = t - win / 2, b = a + win - 1;
      if (a < 0) a = 0;
      if (b > width - 1) b = width - 1;
      w[t] += (float)((pre[b + 1] - pre[a]) / (double)win);
    }
  }
  for (t = 0; t < width; ++t)
    if (!isfinite(w[t]) || w[t] < MINVAL) w[t] = MINVAL;
}

static float spec_taper(int col, int i1, int i2, int i3, int i4) {
  if (col < i1) return 0.f;
  if (col < i2) return (float)((1.0 - cos(M_PI / (i2 - i1) * (col - i1))) / 2.0);
  if (col >= i3 && col < i4) return (float)((1.0 + cos(M_PI / (i4 - i3) * (col - i3))) / 2.0);
  if (col >= i4) return 0.f;
  return 1.f;
}

/* In-place whitening of the frame's nch rows x[ch][seg]. */
static void whiten_frame(const s2s_ctx *c, float *x, cf *X, float *amp, float *ww, double *pre,
                         cf *z, cf *wk) {
  const int seg = c->seg, nb = seg / 2 + 1;
  const int i4 = c->idx[3] < nb ? c->idx[3] : nb;
  int ch, k;
  for (ch = 0; ch < c->nch; ++ch) {
    cf *Xc = X + (long)ch * nb;
    rfft_fwd(c->ps_f, c->fold_s, x + (long)ch * seg, seg, Xc, z, wk);
    for (k = 0; k < nb; ++k) {
      Xc[k].re *= c->dt;
      Xc[k].im *= c->dt;
      if (!finite_cf(Xc[k])) Xc[k].re = Xc[k].im = 0.f;
      amp[(long)ch * nb + k] = sqrtf(Xc[k].re * Xc[k].re + Xc[k].im * Xc[k].im);
    }
  }
  smooth_weights(amp, c->nch, nb, nb, c->whiten_win, ww, pre);
  for (ch = 0; ch < c->nch; ++ch) {
    cf *Xc = X + (long)ch * nb;
    float *y = x + (long)ch * seg;
    const double inv = 1.0 / (double)((float)seg * c->dt);
    for (k = 0; k < nb; ++k) {
      const float d = ww[k];
      cf v = Xc[k];
      if (!isfinite(d) || d < MINVAL) {
        v.re = v.im = 0.f;
      } else {
        v.re /= d;
        v.im /= d;
      }
      {
        const float tp = spec_taper(k, c->idx[0], c->idx[1], c->idx[2], i4);
        v.re *= tp;
        v.im *= tp;
      }
      if (!finite_cf(v)) v.re = v.im = 0.f;
      Xc[k] = v;
    }
    rfft_inv(c->ps_i, c->ifold_s, Xc, y, z, wk);
    for (k = 0; k < seg; ++k) y[k] = (float)((double)y[k] * inv);
  }
}

static void *s2s_worker(void *arg) {
  s2s_ctx *c = (s2s_ctx *)arg;
  const int seg = c->seg, nch = c->nch, nspec = c->nfft_o / 2 + 1;
  const int Mmax = (c->nfft_f > c->nfft_o ? c->nfft_f : c->nfft_o) / 2 + 1;
  float *x = (float *)malloc(sizeof(float) * (size_t)nch * c->nfft_f);
  float *w = (float *)malloc(sizeof(float) * (size_t)(seg + 1));
  float *amp = (float *)malloc(sizeof(float) * (size_t)nch * (seg / 2 + 1));
  float *ww = (float *)malloc(sizeof(float) * (size_t)(seg / 2 + 1));
  double *pre = (double *)malloc(sizeof(double) * (size_t)(seg + 2));
  cf *X = (cf *)malloc(sizeof(cf) * (size_t)nch * Mmax);
  cf *z = (cf *)malloc(sizeof(cf) * (size_t)Mmax);
  cf *wk = (cf *)malloc(sizeof(cf) * (size_t)Mmax);
  float *full = (float *)malloc(sizeof(float) * (size_t)c->nfft_f);
  float *ab = (float *)malloc(sizeof(float) * (size_t)nch * seg);
  const int nbf = c->nfft_f / 2 + 1;
  const double inv_f = 1.0 / (double)((float)c->nfft_f * c->dt);
  const int groups = c->nfiles / nch;
  long fr;
  for (fr = c->frame_begin; fr < c->frame_end; ++fr) {
    const int step = (int)(fr / groups), g = (int)(fr % groups);
    const long off = (long)c->steps[step] * c->shift;
    int ch, i, k;
    for (ch = 0; ch < nch; ++ch) {
      const float *src = c->traces + (long)(g * nch + ch) * c->npts;
      float *y = x + (long)ch * seg;
      double s = 0.0, s1 = 0.0, si = 0.0;
      for (i = 0; i < seg; ++i) {
        float v = off + i < c->npts ? src[off + i] : 0.f;
        if (!isfinite(v)) v = 0.f;
        y[i] = v;
        s += v;
      }
      {
        const double mean = s / seg;
        for (i = 0; i < seg; ++i) {
          const float v = (float)((double)y[i] - mean);
          y[i] = v;
          s1 += v;
          si += (double)v * i;
        }
      }
      {
        const double W = seg, a12 = 0.5 * W * (W - 1), a11 = a12 * (2 * W - 1) / 3., a22 = W;
        const double det = a11 * a22 - a12 * a12;
        const double ca = (a22 * si - a12 * s1) / det, cb = (a11 * s1 - a12 * si) / det;
        const int T = c->taper;
        for (i = 0; i < seg; ++i) {
          const float v = (float)((double)y[i] - (ca * i + cb));
          float f = 1.f;
          if (i < T)
            f = (float)(0.5 * (1.0 - cos(M_PI * i / T)));
          else if (i >= seg - T)
            f = (float)(0.5 * (1.0 - cos(M_PI * (seg - i) / T)));
          y[i] = v * f;
        }
      }
      /* prefilter: R2C(nfft_f) x dt, sanitise, x |H0|^2, C2R, x 1/(nfft_f dt) */
      rfft_fwd(c->pf_f, c->fold_f, y, seg, X, z, wk);
      for (k = 0; k < nbf; ++k) {
        X[k].re *= c->dt;
        X[k].im *= c->dt;
        if (!finite_cf(X[k])) X[k].re = X[k].im = 0.f;
        X[k].re *= c->resp0[k];
        X[k].im *= c->resp0[k];
      }
      rfft_inv(c->pf_i, c->ifold_f, X, full, z, wk);
      for (i = 0; i < seg; ++i) y[i] = (float)((double)full[i] * inv_f);
    }
    if (c->whiten == 1 || c->whiten == 3) whiten_frame(c, x, X, amp, ww, pre, z, wk);
    if (c->normalize == 3) {
      for (i = 0; i < nch * seg; ++i) ab[i] = fabsf(x[i]);
      smooth_weights(ab, nch, seg, seg, c->runabs_win, w, pre);
      for (ch = 0; ch < nch; ++ch)
        for (i = 0; i < seg; ++i) {
          float v = x[(long)ch * seg + i];
          const float d = w[i];
          v = (!isfinite(d) || d < MINVAL) ? 0.f : v / d;
          if (!isfinite(v)) v = 0.f;
          v = v > MAXVAL ? MAXVAL : (v < -MAXVAL ? -MAXVAL : v);
          x[(long)ch * seg + i] = v;
        }
    } else if (c->normalize == 2) {
      for (i = 0; i < nch * seg; ++i) x[i] = x[i] > 0.f ? 1.f : (x[i] < 0.f ? -1.f : 0.f);
    }
    if (c->whiten == 2 || c->whiten == 3) whiten_frame(c, x, X, amp, ww, pre, z, wk);
    for (ch = 0; ch < nch; ++ch) {
      cf *o = (cf *)c->out + ((long)step * c->nfiles + g * nch + ch) * nspec;
      rfft_fwd(c->po_f, c->fold_o, x + (long)ch * seg, seg, o, z, wk);
      for (k = 0; k < nspec; ++k) {
        o[k].re *= c->dt;
        o[k].im *= c->dt;
        if (!finite_cf(o[k])) o[k].re = o[k].im = 0.f;
      }
    }
  }
  free(x);
  free(full);
  free(ab);
  free(w);
  free(amp);
  free(ww);
  free(pre);
  free(X);
  free(z);
  free(wk);
  return NULL;
}

/* traces [nfiles][npts] (channels of a group adjacent); steps: window
 * indices; out [nsteps][nfiles][nfft_o/2+1] complex64.  normalize 0 none,
 * 2 one-bit, 3 RUN-ABS; whiten 0 none, 1 before, 2 after, 3 both.
 * Returns 0, or -1 for an unsupported mode or length. */
int cpu_s2s_process(const float *traces, int nfiles, int npts, const int *steps, int nsteps,
                    int seg, int shift, int nfft_f, int nfft_o, int nch, float dt, float freq_low,
                    const int *idx, int whiten, int normalize, const float *resp0, float *out,
                    int nthreads) {
  s2s_ctx base;
  fftplan *pl;
  pthread_t th[256];
  s2s_ctx *ctx;
  long frames;
  int t;
  if (normalize == 1 || nch < 1 || nfiles % nch) return -1;
  memset(&base, 0, sizeof(base));
  base.npts = npts;
  base.seg = seg;
  base.shift = shift;
  base.nfft_f = nfft_f;
  base.nfft_o = nfft_o;
  base.nch = nch;
  base.whiten = whiten;
  base.normalize = normalize;
  base.dt = dt;
  /* (size_t)(2.0f*(1.0f/freq_low)/delta) (preprocess.cu:91) */
  base.taper = (int)(2.0f * (1.0f / freq_low) / dt);
  {
    const float fra = freq_low * 0.667f;
    base.runabs_win = 2 * (int)(1.0 / (double)(fra * dt)) + 1;
  }
  base.whiten_win = (int)(0.02 * seg * (double)dt);
  if (base.whiten_win < 1) base.whiten_win = 1;
  memcpy(base.idx, idx, sizeof(base.idx));
  base.traces = traces;
  base.nfiles = nfiles;
  base.steps = steps;
  base.nsteps = nsteps;
  base.resp0 = resp0;
  base.out = out;
  pl = (fftplan *)calloc(5, sizeof(fftplan));
  if (fft_plan_init(&pl[0], nfft_f / 2, -1) || fft_plan_init(&pl[1], nfft_f / 2, 1) ||
      fft_plan_init(&pl[2], seg / 2, -1) || fft_plan_init(&pl[3], seg / 2, 1) ||
      fft_plan_init(&pl[4], nfft_o / 2, -1)) {
    for (t = 0; t < 5; ++t) free(pl[t].tw);
    free(pl);
    return -1;
  }
  base.pf_f = &pl[0];
  base.pf_i = &pl[1];
  base.ps_f = &pl[2];
  base.ps_i = &pl[3];
  base.po_f = &pl[4];
  base.fold_f = fold_table(nfft_f, -1);
  base.ifold_f = fold_table(nfft_f, 1);
  base.fold_s = fold_table(seg, -1);
  base.ifold_s = fold_table(seg, 1);
  base.fold_o = fold_table(nfft_o, -1);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  frames = (long)nsteps * (nfiles / nch);
  ctx = (s2s_ctx *)calloc((size_t)nthreads, sizeof(s2s_ctx));
  for (t = 0; t < nthreads; ++t) {
    ctx[t] = base;
    ctx[t].frame_begin = frames * t / nthreads;
    ctx[t].frame_end = frames * (t + 1) / nthreads;
    pthread_create(&th[t], NULL, s2s_worker, &ctx[t]);
  }
  for (t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  for (t = 0; t < 5; ++t) fft_plan_free(&pl[t]);
  free(pl);
  free(ctx);
  free(base.fold_f);
  free(base.ifold_f);
  free(base.fold_s);
  free(base.ifold_s);
  free(base.fold_o);
  return 0;
}

/* ------------------------------------------------------------------ stacks */
typedef struct {
  const float *pre, *w, *lin;
  const long long *off;
  int n, kind;
  float *out;
  const fftplan *pf, *pi;
  long begin, end;
} stk_ctx;

/* Analytic spectrum of a real row: bins 1..(n+1)/2-1 doubled, the rest 0. */
static void analytic(const fftplan *pf, const float *x, int n, cf *A, cf *wk) {
  const int half = (n + 1) / 2;
  int k;
  for (k = 0; k < n; ++k) {
    A[k].re = x[k];
    A[k].im = 0.f;
  }
  fft_exec(pf, A, wk);
  for (k = 0; k < n; ++k) {
    if (k >= 1 && k < half) {
      A[k].re *= 2.f;
      A[k].im *= 2.f;
    } else {
      A[k].re = A[k].im = 0.f;
    }
  }
}

static void *pws_worker(void *arg) {
  stk_ctx *c = (stk_ctx *)arg;
  const int n = c->n;
  cf *a = (cf *)malloc(sizeof(cf) * (size_t)n), *wk = (cf *)malloc(sizeof(cf) * (size_t)n);
  cf *W = (cf *)malloc(sizeof(cf) * (size_t)n);
  const float invn = 1.f / (float)n;
  long p;
  for (p = c->begin; p < c->end; ++p) {
    long long g;
    int t;
    for (t = 0; t < n; ++t) W[t].re = W[t].im = 0.f;
    for (g = c->off[p]; g < c->off[p + 1]; ++g) {
      analytic(c->pf, c->pre + g * n, n, a, wk);
      fft_exec(c->pi, a, wk);
      for (t = 0; t < n; ++t) {
        const float re = a[t].re * invn, im = a[t].im * invn;
        float mod = sqrtf(re * re + im * im);
        mod = mod > 1e-7f ? mod : 1e-7f;
        W[t].re += (re / mod) * c->w[g];
        W[t].im += (im / mod) * c->w[g];
      }
    }
    for (t = 0; t < n; ++t)
      c->out[p * n + t] = c->lin[p * n + t] * sqrtf(W[t].re * W[t].re + W[t].im * W[t].im);
  }
  free(a);
  free(wk);
  free(W);
  return NULL;
}

/* G(c, f) of gaussianModulateSub, scale 0.1 */
static float tf_gauss(int n, int c, int f) {
  const int half = (n + 1) / 2;
  int cm;
  if (f == 0) return 1.f;
  if (c == 0) return 0.f;
  cm = c < half ? c : n - c;
  if (cm >= half) return 0.f;
  return expf(-2.0f * (float)M_PI * (float)M_PI * (float)cm * (float)cm * 0.1f /
              ((float)f * (float)f));
}

static void *tfpws_worker(void *arg) {
  stk_ctx *c = (stk_ctx *)arg;
  const int n = c->n, nfreq = n / 2 + 1;
  cf *wk = (cf *)malloc(sizeof(cf) * (size_t)n), *Al = (cf *)malloc(sizeof(cf) * (size_t)n);
  cf *st = (cf *)malloc(sizeof(cf) * (size_t)n), *W = (cf *)malloc(sizeof(cf) * (size_t)n);
  cf *spec = (cf *)malloc(sizeof(cf) * (size_t)n);
  float *gs = (float *)malloc(sizeof(float) * (size_t)n);
  const float nf = (float)n;
  long p;
  for (p = c->begin; p < c->end; ++p) {
    const long long g0 = c->off[p], g1 = c->off[p + 1];
    cf *Ag = (cf *)malloc(sizeof(cf) * (size_t)n * (size_t)(g1 - g0));
    long long g;
    int f, t;
    for (g = g0; g < g1; ++g) analytic(c->pf, c->pre + g * n, n, Ag + (g - g0) * n, wk);
    analytic(c->pf, c->lin + p * n, n, Al, wk);
    for (t = 0; t < n; ++t) spec[t].re = spec[t].im = 0.f;
    for (f = 0; f < nfreq; ++f) {
      float sx = 0.f, sy = 0.f;
      for (t = 0; t < n; ++t) gs[t] = tf_gauss(n, t, f);
      for (t = 0; t < n; ++t) W[t].re = W[t].im = 0.f;
      for (g = g0; g < g1; ++g) {
        const cf *A = Ag + (g - g0) * n;
        for (t = 0; t < n; ++t) {
          int s = f + t;
          if (s >= n) s -= n;
          st[t].re = A[s].re * gs[t];
          st[t].im = A[s].im * gs[t];
        }
        fft_exec(c->pi, st, wk);
        for (t = 0; t < n; ++t) {
          float fa = sqrtf(st[t].re * st[t].re + st[t].im * st[t].im);
          if (fa <= 1e-7f) fa = 1.f;
          W[t].re += c->w[g] * st[t].re / (fa * nf);
          W[t].im += c->w[g] * st[t].im / (fa * nf);
        }
      }
      for (t = 0; t < n; ++t) {
        int s = f + t;
        if (s >= n) s -= n;
        st[t].re = Al[s].re * gs[t];
        st[t].im = Al[s].im * gs[t];
      }
      fft_exec(c->pi, st, wk);
      for (t = 0; t < n; ++t) {
        const float mag = sqrtf(W[t].re * W[t].re + W[t].im * W[t].im);
        sx += st[t].re * mag;
        sy += st[t].im * mag;
      }
      spec[f].re = sx;
      spec[f].im = sy;
    }
    fft_exec(c->pi, spec, wk);
    for (t = 0; t < n; ++t) c->out[p * n + t] = spec[t].re / nf;
    free(Ag);
  }
  free(wk);
  free(Al);
  free(st);
  free(W);
  free(spec);
  free(gs);
  return NULL;
}

static int run_stack(const float *pre, const float *w, const long long *off, const float *lin,
                     int npairs, int n, float *out, int nthreads, int kind) {
  fftplan *pl = (fftplan *)calloc(2, sizeof(fftplan));
  pthread_t th[256];
  stk_ctx ctx[256];
  int t;
  if (fft_plan_init(&pl[0], n, -1) || fft_plan_init(&pl[1], n, 1)) {
    free(pl[0].tw);
    free(pl[1].tw);
    free(pl);
    return -1;
  }
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  for (t = 0; t < nthreads; ++t) {
    ctx[t].pre = pre;
    ctx[t].w = w;
    ctx[t].off = off;
    ctx[t].lin = lin;
    ctx[t].n = n;
    ctx[t].kind = kind;
    ctx[t].out = out;
    ctx[t].pf = &pl[0];
    ctx[t].pi = &pl[1];
    ctx[t].begin = (long)npairs * t / nthreads;
    ctx[t].end = (long)npairs * (t + 1) / nthreads;
    pthread_create(&th[t], NULL, kind ? tfpws_worker : pws_worker, &ctx[t]);
  }
  for (t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  fft_plan_free(&pl[0]);
  fft_plan_free(&pl[1]);
  free(pl);
  return 0;
}

/* PWS (power 1): prestack rows are group means, w = n_g/T, off [npairs+1]. */
int cpu_pws(const float *pre, const float *w, const long long *off, const float *lin, int npairs,
            int n, float *out, int nthreads) {
  return run_stack(pre, w, off, lin, npairs, n, out, nthreads, 0);
}

/* TF-PWS (power 1, every voice, no band blend): prestack rows are group sums. */
int cpu_tfpws(const float *pre, const float *w, const long long *off, const float *lin,
              int npairs, int n, float *out, int nthreads) {
  return run_stack(pre, w, off, lin, npairs, n, out, nthreads, 1);
}

typedef struct {
  const float *rec;
  const long long *off;
  int n;
  float *out;
  long begin, end;
} lin_ctx;

static void *lin_worker(void *arg) {
  lin_ctx *c = (lin_ctx *)arg;
  long p;
  for (p = c->begin; p < c->end; ++p) {
    const long long r0 = c->off[p], r1 = c->off[p + 1];
    int j;
    for (j = 0; j < c->n; ++j) {
      double acc = 0.0;
      long long r;
      for (r = r0; r < r1; ++r) acc += (double)c->rec[r * c->n + j];
      c->out[p * c->n + j] = (float)(acc / (double)(r1 - r0));
    }
  }
  return NULL;
}

/* float64 running sum of each pair's rows in row order / count, one rounding. */
int cpu_linear_stack(const float *rec, const long long *off, int npairs, int n, float *out,
                     int nthreads) {
  pthread_t th[256];
  lin_ctx ctx[256];
  int t;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  for (t = 0; t < nthreads; ++t) {
    ctx[t].rec = rec;
    ctx[t].off = off;
    ctx[t].n = n;
    ctx[t].out = out;
    ctx[t].begin = (long)npairs * t / nthreads;
    ctx[t].end = (long)npairs * (t + 1) / nthreads;
    pthread_create(&th[t], NULL, lin_worker, &ctx[t]);
  }
  for (t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return 0;
}
