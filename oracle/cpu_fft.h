/*
 * Mixed-radix Stockham FFT for the C restatements in oracle/ -- TEST AND
 * BASELINE INFRASTRUCTURE ONLY.
 *
 * Radices 4, 2, 3, 5, 7, 23 and 29 (every length the stages use: the window,
 * prefilter and output lengths 36000, 36288, 37044, their halves, and the
 * correlation length 2001 = 3 x 23 x 29).  Odd radices are evaluated directly
 * from a table of their roots of unity (float), twiddles come from FP64-
 * computed tables, data are float.  Unnormalised: sign -1 forward, +1 inverse.
 */
#ifndef ORACLE_CPU_FFT_H
#define ORACLE_CPU_FFT_H

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  float re, im;
} cf;

#define CPU_FFT_MAXR 29

typedef struct {
  int n, npass, sign;
  int radix[40];
  int ns[40];
  int tw_off[40];
  cf *tw;                           /* per pass: ns*(r-1) twiddles exp(sign 2 pi i m k/(ns r)) */
  float rc[CPU_FFT_MAXR + 1][CPU_FFT_MAXR]; /* roots of each odd radix: cos, sin of 2 pi m/r */
  float rs[CPU_FFT_MAXR + 1][CPU_FFT_MAXR];
} fftplan;

static inline int fft_plan_init(fftplan *p, int n, int sign) {
  static const int prefs[] = {4, 2, 3, 5, 7, 23, 29};
  int m = n, cnt = 0, i, total = 0, ns = 1;
  memset(p, 0, sizeof(*p));
  p->n = n;
  p->sign = sign;
  for (i = 0; i < 7; ++i)
    while (m % prefs[i] == 0 && m > 1) {
      p->radix[cnt++] = prefs[i];
      m /= prefs[i];
    }
  if (m != 1 || cnt > 40) return -1;
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
  for (i = 3; i <= CPU_FFT_MAXR; i += 2) {
    int q;
    for (q = 0; q < i; ++q) {
      double a = sign * 2.0 * M_PI * (double)q / i;
      p->rc[i][q] = (float)cos(a);
      p->rs[i][q] = (float)sin(a);
    }
  }
  return 0;
}

static inline void fft_plan_free(fftplan *p) { free(p->tw); }

/* Stockham autosort FFT of x (n values); work: scratch of n. */
static inline void fft_exec(const fftplan *p, cf *x, cf *work) {
  int n = p->n, pi, j;
  cf *in = x, *out = work;
  for (pi = 0; pi < p->npass; ++pi) {
    const int r = p->radix[pi], ns = p->ns[pi], nb = n / r;
    const cf *tw = p->tw + p->tw_off[pi];
    for (j = 0; j < nb; ++j) {
      const int k = j % ns;
      cf a[CPU_FFT_MAXR], b[CPU_FFT_MAXR];
      int m, q;
      for (m = 0; m < r; ++m) a[m] = in[j + m * nb];
      if (ns > 1)
        for (m = 1; m < r; ++m) {
          cf w = tw[k * (r - 1) + m - 1], v = a[m];
          a[m].re = v.re * w.re - v.im * w.im;
          a[m].im = v.re * w.im + v.im * w.re;
        }
      if (r == 2) {
        b[0].re = a[0].re + a[1].re; b[0].im = a[0].im + a[1].im;
        b[1].re = a[0].re - a[1].re; b[1].im = a[0].im - a[1].im;
      } else if (r == 4) {
        const float sg = (float)p->sign;
        cf s02 = {a[0].re + a[2].re, a[0].im + a[2].im}, d02 = {a[0].re - a[2].re, a[0].im - a[2].im};
        cf s13 = {a[1].re + a[3].re, a[1].im + a[3].im}, d13 = {a[1].re - a[3].re, a[1].im - a[3].im};
        cf id13 = {-sg * d13.im, sg * d13.re};
        b[0].re = s02.re + s13.re; b[0].im = s02.im + s13.im;
        b[2].re = s02.re - s13.re; b[2].im = s02.im - s13.im;
        b[1].re = d02.re + id13.re; b[1].im = d02.im + id13.im;
        b[3].re = d02.re - id13.re; b[3].im = d02.im - id13.im;
      } else {
        const float *rc = p->rc[r], *rs = p->rs[r];
        for (q = 0; q < r; ++q) {
          float sr = 0.f, si = 0.f;
          int e = 0;
          for (m = 0; m < r; ++m) {
            sr += a[m].re * rc[e] - a[m].im * rs[e];
            si += a[m].re * rs[e] + a[m].im * rc[e];
            e += q;
            if (e >= r) e -= r;
          }
          b[q].re = sr;
          b[q].im = si;
        }
      }
      {
        cf *dst = out + (j / ns) * ns * r + k;
        for (q = 0; q < r; ++q) dst[q * ns] = b[q];
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

/* Real forward transform of x[0..n) (n even, zero past len): bins 0..n/2 of
 * X; z, w: scratch of n/2; half: plan of n/2 (sign -1); fold[k] =
 * exp(-2 pi i k / n), k <= n/2. */
static inline void rfft_fwd(const fftplan *half, const cf *fold, const float *x, int len, cf *X, cf *z,
                     cf *w) {
  const int M = half->n;
  int k;
  for (k = 0; k < M; ++k) {
    z[k].re = 2 * k < len ? x[2 * k] : 0.f;
    z[k].im = 2 * k + 1 < len ? x[2 * k + 1] : 0.f;
  }
  fft_exec(half, z, w);
  for (k = 0; k <= M; ++k) {
    cf a = z[k == M ? 0 : k], b = z[k == 0 ? 0 : M - k];
    float er = 0.5f * (a.re + b.re), ei = 0.5f * (a.im - b.im);
    float dr = 0.5f * (a.re - b.re), di = 0.5f * (a.im + b.im);
    /* X = e - i w d  with d = (a - conj b)/2 */
    float wr = fold[k].re * dr - fold[k].im * di, wi = fold[k].re * di + fold[k].im * dr;
    X[k].re = er + wi;
    X[k].im = ei - wr;
  }
}

/* Real inverse transform (unnormalised) of bins 0..n/2 into y[0..n): z, w
 * scratch of n/2; half: plan of n/2 (sign +1); ifold[k] = exp(+2 pi i k / n). */
static inline void rfft_inv(const fftplan *half, const cf *ifold, const cf *X, float *y, cf *z, cf *w) {
  const int M = half->n;
  int k;
  for (k = 0; k < M; ++k) {
    cf a = X[k], b = X[M - k];
    float sr, si, dr, di;
    if (k == 0) {
      a.im = 0.f;
      b.im = 0.f;
    }
    b.im = -b.im;
    sr = a.re + b.re;
    si = a.im + b.im;
    dr = a.re - b.re;
    di = a.im - b.im;
    {
      float wr = ifold[k].re * dr - ifold[k].im * di, wi = ifold[k].re * di + ifold[k].im * dr;
      z[k].re = sr - wi;
      z[k].im = si + wr;
    }
  }
  fft_exec(half, z, w);
  for (k = 0; k < M; ++k) {
    y[2 * k] = z[k].re;
    y[2 * k + 1] = z[k].im;
  }
}

static inline cf *fold_table(int n, int sign) {
  const int M = n / 2;
  cf *f = (cf *)malloc(sizeof(cf) * (size_t)(M + 1));
  int k;
  for (k = 0; k <= M; ++k) {
    double a = sign * 2.0 * M_PI * (double)k / n;
    f[k].re = (float)cos(a);
    f[k].im = (float)sin(a);
  }
  return f;
}

#endif /* ORACLE_CPU_FFT_H */
