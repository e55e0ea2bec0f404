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
 * The C2R runs as one complex FFT of length nfft/2 (half-length fold) with a
 * mixed-radix (2,3,4,5,7) Stockham FFT written here; lag m reads sample
 * (m - H) mod nfft of the unshifted C2R, which equals the reference's
 * (-1)^k shift.  Pairs are split over pthreads in contiguous blocks.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  float re, im;
} cf;

typedef struct {
  int n, npass;
  int radix[40];
  int ns[40];
  int tw_off[40];
  cf *tw; /* per pass: ns*(r-1) twiddles exp(sign 2 pi i m k/(ns r)) */
  int sign;
} fftplan;

static int plan_init(fftplan *p, int n, int sign) {
  static const int prefs[] = {4, 2, 3, 5, 7};
  int m = n, cnt = 0, i, total = 0, ns = 1;
  memset(p, 0, sizeof(*p));
  p->n = n;
  p->sign = sign;
  for (i = 0; i < 5; ++i)
    while (m % prefs[i] == 0 && m > 1) {
      p->radix[cnt++] = prefs[i];
      m /= prefs[i];
    }
  if (m != 1) return -1;
  p->npass = cnt;
  for (i = 0; i < cnt; ++i) {
    p->ns[i] = ns;
    p->tw_off[i] = total;
    total += ns * (p->radix[i] - 1);
    ns *= p->radix[i];
  }
  p->tw = (cf *)malloc(sizeof(cf) * (size_t)(total > 0 ? total : 1));
  for (i = 0; i < cnt; ++i) {
    int r = p->radix[i], k, q;
    for (k = 0; k < p->ns[i]; ++k)
      for (q = 1; q < r; ++q) {
        double a = sign * 2.0 * M_PI * (double)((long long)q * k % ((long long)p->ns[i] * r)) /
                   ((double)p->ns[i] * r);
        p->tw[p->tw_off[i] + k * (r - 1) + q - 1].re = (float)cos(a);
        p->tw[p->tw_off[i] + k * (r - 1) + q - 1].im = (float)sin(a);
      }
  }
  return 0;
}

static void plan_free(fftplan *p) { free(p->tw); }

/* b_q = sum_m a_m exp(sign 2 pi i m q / r), small r */
static void dft_small(cf *a, int r, int sign) {
  cf b[8];
  int q, m;
  if (r == 2) {
    cf t = a[0];
    a[0].re = t.re + a[1].re; a[0].im = t.im + a[1].im;
    a[1].re = t.re - a[1].re; a[1].im = t.im - a[1].im;
    return;
  }
  if (r == 4) {
    cf s02 = {a[0].re + a[2].re, a[0].im + a[2].im}, d02 = {a[0].re - a[2].re, a[0].im - a[2].im};
    cf s13 = {a[1].re + a[3].re, a[1].im + a[3].im}, d13 = {a[1].re - a[3].re, a[1].im - a[3].im};
    cf id13 = {-sign * d13.im, sign * d13.re};
    a[0].re = s02.re + s13.re; a[0].im = s02.im + s13.im;
    a[2].re = s02.re - s13.re; a[2].im = s02.im - s13.im;
    a[1].re = d02.re + id13.re; a[1].im = d02.im + id13.im;
    a[3].re = d02.re - id13.re; a[3].im = d02.im - id13.im;
    return;
  }
  for (q = 0; q < r; ++q) {
    double sr = 0, si = 0;
    for (m = 0; m < r; ++m) {
      double ang = sign * 2.0 * M_PI * (double)((m * q) % r) / r;
      double c = cos(ang), s = sin(ang);
      sr += a[m].re * c - a[m].im * s;
      si += a[m].re * s + a[m].im * c;
    }
    b[q].re = (float)sr;
    b[q].im = (float)si;
  }
  for (q = 0; q < r; ++q) a[q] = b[q];
}

/* Precomputed small-DFT roots for r in {3,5,7} (float). */
typedef struct {
  float c[8][8], s[8][8];
} roots_t;

static void roots_init(roots_t *rt, int r, int sign) {
  int q, m;
  for (q = 0; q < r; ++q)
    for (m = 0; m < r; ++m) {
      double ang = sign * 2.0 * M_PI * (double)((m * q) % r) / r;
      rt->c[q][m] = (float)cos(ang);
      rt->s[q][m] = (float)sin(ang);
    }
}

/* Stockham autosort FFT; result ends in x (work is scratch of n). */
static void fft_exec(const fftplan *p, const roots_t *rts, cf *x, cf *work) {
  int n = p->n, pi, j;
  cf *in = x, *out = work;
  for (pi = 0; pi < p->npass; ++pi) {
    const int r = p->radix[pi], ns = p->ns[pi], nb = n / r;
    const cf *tw = p->tw + p->tw_off[pi];
    for (j = 0; j < nb; ++j) {
      const int k = j % ns;
      cf a[8];
      int m, q;
      for (m = 0; m < r; ++m) a[m] = in[j + m * nb];
      if (ns > 1)
        for (m = 1; m < r; ++m) {
          cf w = tw[k * (r - 1) + m - 1], v = a[m];
          a[m].re = v.re * w.re - v.im * w.im;
          a[m].im = v.re * w.im + v.im * w.re;
        }
      if (r == 2 || r == 4) {
        dft_small(a, r, p->sign);
      } else {
        cf b[8];
        const roots_t *rt = &rts[r];
        for (q = 0; q < r; ++q) {
          float sr = 0.f, si = 0.f;
          for (m = 0; m < r; ++m) {
            sr += a[m].re * rt->c[q][m] - a[m].im * rt->s[q][m];
            si += a[m].re * rt->s[q][m] + a[m].im * rt->c[q][m];
          }
          b[q].re = sr;
          b[q].im = si;
        }
        for (q = 0; q < r; ++q) a[q] = b[q];
      }
      {
        cf *dst = out + (j / ns) * ns * r + k;
        for (q = 0; q < r; ++q) dst[q * ns] = a[q];
      }
    }
    {
      cf *t = in;
      in = out;
      out = t;
    }
  }
  if (in != x) memcpy(x, in, sizeof(cf) * (size_t)n);
}

typedef struct {
  const cf *spec;
  int nstep, nsta, nspec, half_cc;
  float dt;
  const int *src, *rec;
  long begin, end;
  float *out;
  const fftplan *plan;
  const roots_t *rts;
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
    fft_exec(jb->plan, jb->rts, Z, W);
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
  roots_t rts[8];
  const int M = nspec - 1;
  cf *fold;
  pthread_t th[256];
  job_t jobs[256];
  int t, k;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (plan_init(&plan, M, +1) != 0) return -1;
  for (k = 3; k < 8; ++k) roots_init(&rts[k], k, +1);
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
    jobs[t].rts = rts;
    jobs[t].fold = fold;
    pthread_create(&th[t], NULL, worker, &jobs[t]);
  }
  for (t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(fold);
  plan_free(&plan);
  return 0;
}
