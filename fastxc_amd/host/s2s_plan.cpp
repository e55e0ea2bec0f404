#include "s2s_plan.hpp"

#include "fxc_host.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace fxh {

namespace {
// five numbers separated by blanks or tabs
bool parse_coeffs(const char *line, double *out) {
  const char *p = line;
  for (int i = 0; i < 5; ++i) {
    char *end = nullptr;
    out[i] = strtod(p, &end);
    if (end == p) return false;
    p = end;
  }
  return true;
}
}  // namespace

std::vector<Butterworth> read_filters(const std::string &path) {
  std::vector<Butterworth> out;
  FILE *fp = fopen(path.c_str(), "r");
  if (!fp) {
    FXH_LOG_ERROR("filter_file_open_failed", "path=\"%s\"", path.c_str());
    return out;
  }
  char line[1024];
  int state = 0;
  Butterworth cur;
  while (fgets(line, sizeof(line), fp)) {
    if (line[0] == '#') {
      if (sscanf(line, "# %f/%f", &cur.freq_low, &cur.freq_high) != 2) continue;
      state = 1;
      continue;
    }
    if (state == 1) {
      if (!parse_coeffs(line, cur.b)) continue;
      state = 2;
    } else if (state == 2) {
      if (!parse_coeffs(line, cur.a)) {
        state = 0;
        continue;
      }
      out.push_back(cur);
      cur = Butterworth();
      state = 0;
    }
  }
  fclose(fp);
  return out;
}

bool is_smooth_2357(int v) {
  if (v < 1) return false;
  for (int p : {2, 3, 5, 7})
    while (v > 1 && v % p == 0) v /= p;
  return v == 1;
}

int next_smooth_2357(int n) {
  n = std::max(n, 1);
  while (!is_smooth_2357(n)) ++n;
  return n;
}

int next_even_smooth_2357(int n) {
  int v = next_smooth_2357(n);
  while (v % 2) v = next_smooth_2357(v + 1);
  return v;
}

int previous_smooth_2357(int n) {
  for (int v = n - 1; v > 0; --v)
    if (is_smooth_2357(v)) return v;
  return 0;
}

int segment_length_from_seconds(float seconds, int max_samples, float delta) {
  const int req = (seconds <= 0.f || delta <= 0.f) ? 0 : (int)llround((double)seconds / (double)delta);
  int seg = next_smooth_2357(req);
  if (seg > max_samples) seg = previous_smooth_2357(max_samples);
  return seg;
}

// Length that holds all but 1e-6 of the single-pass impulse response energy.
int impulse_length(const Butterworth &f, int max_samples) {
  const double a0 = f.a[0];
  if (std::fabs(a0) < 1e-30) return max_samples;
  double x[5] = {0, 0, 0, 0, 0}, y[5] = {0, 0, 0, 0, 0};
  std::vector<double> resp((size_t)max_samples);
  double total = 0.0;
  for (int n = 0; n < max_samples; ++n) {
    for (int j = 4; j > 0; --j) {
      x[j] = x[j - 1];
      y[j] = y[j - 1];
    }
    x[0] = n == 0 ? 1.0 : 0.0;
    double v = 0.0;
    for (int j = 0; j < 5; ++j) v += f.b[j] * x[j];
    for (int j = 1; j < 5; ++j) v -= f.a[j] * y[j];
    v /= a0;
    y[0] = v;
    resp[n] = v;
    total += v * v;
  }
  if (total <= 0.0) return 1;
  double tail = 0.0;
  for (int n = max_samples - 1; n >= 0; --n) {
    tail += resp[n] * resp[n];
    if (tail > total * 1e-6) return n + 1;
  }
  return 1;
}

int filter_padding(const std::vector<Butterworth> &fs, int max_padding) {
  int m = 1;
  for (const auto &f : fs) m = std::max(m, impulse_length(f, max_padding));
  return std::min(std::max(2 * m, 1), max_padding);
}

std::vector<float> power_response(const Butterworth &f, int nfft) {
  const int half = nfft / 2;
  std::vector<float> r((size_t)half + 1, 0.f);
  if (half <= 0) return r;
  for (int i = 0; i <= half; ++i) {
    const double w = M_PI * (double)i / (double)half;
    double bx = 0, by = 0, ax = 0, ay = 0;
    for (int j = 0; j < 5; ++j) {
      const double c = std::cos(-w * j), s = std::sin(-w * j);
      bx += f.b[j] * c;
      by += f.b[j] * s;
      ax += f.a[j] * c;
      ay += f.a[j] * s;
    }
    const double num = bx * bx + by * by, den = ax * ax + ay * ay;
    double p = den <= DBL_MIN ? 0.0 : num / den;
    if (!std::isfinite(p)) p = 0.0;
    if (p > (double)FLT_MAX) p = (double)FLT_MAX;
    r[i] = (float)p;
  }
  return r;
}

bool derive_lengths(float sac_len, float delta, float win_len, float shift_len, float max_lag,
                    const std::vector<Butterworth> &filters, S2sLengths *o) {
  const double target = (double)sac_len / (double)delta;
  if (!(delta > 0.f) || !std::isfinite(target) || target < 1.0 || target > 2147483647.0) {
    FXH_LOG_ERROR("invalid_sac_length", "sac_len_sec=%.9g delta=%.9g", (double)sac_len,
                  (double)delta);
    return false;
  }
  o->npts = (int)llround(target);
  o->seg = segment_length_from_seconds(win_len, o->npts, delta);
  o->shift = (int)llround((double)shift_len / (double)delta);
  if (o->shift < 1) o->shift = o->seg;
  o->nstep = o->seg <= o->npts ? std::max((o->npts - o->seg) / o->shift + 1, 1) : 1;
  const int max_out = 2 * o->seg;
  int keep = (int)llround((double)max_lag / (double)delta);
  keep = std::min(keep, o->seg);
  o->nfft_out = std::min(next_even_smooth_2357(o->seg + keep), max_out);
  o->nspec = o->nfft_out / 2 + 1;
  o->df_out = 1.0f / ((float)o->nfft_out * delta);
  if (filters.empty()) {
    FXH_LOG_ERROR("filter_file_empty", "bands=0");
    return false;
  }
  const float flo = filters[0].freq_low, fhi = filters[0].freq_high;
  if (flo <= 0.f || fhi <= flo) {
    FXH_LOG_ERROR("filter_band_invalid", "freq_low=%.9g freq_high=%.9g", (double)flo, (double)fhi);
    return false;
  }
  const float seg_df = 1.0f / ((float)o->seg * delta);
  o->idx[0] = (int)(flo * 0.667f / seg_df);
  o->idx[1] = (int)(flo / seg_df);
  o->idx[2] = (int)(fhi / seg_df);
  o->idx[3] = (int)(fhi * 1.333f / seg_df);
  const int pad = filter_padding(filters, o->seg);
  o->nfft_filter = std::min(next_even_smooth_2357(o->seg + pad), max_out);
  return true;
}

}  // namespace fxh
