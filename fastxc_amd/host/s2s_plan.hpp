// sac2spec plan derivation on the host: window and FFT lengths, the
// whitening taper bins and the Butterworth power responses.
//
// Follows native/sac2spec/cuda.main.cu:190-296 (integer decisions taken in
// float32 as the reference takes them), fft_length.c:20-109
// (2,3,5,7-smooth lengths) and core/filter_response.c:10-420 (filter.txt
// parsing, impulse-length padding, |B|^2/|A|^2 in double stored as float).
#pragma once

#include <string>
#include <vector>

namespace fxh {

struct Butterworth {
  float freq_low = 0.f, freq_high = 0.f;
  double b[5] = {0, 0, 0, 0, 0}, a[5] = {0, 0, 0, 0, 0};
};

// "# lo/hi" line, then a line of 5 b and a line of 5 a coefficients per band.
std::vector<Butterworth> read_filters(const std::string &path);

bool is_smooth_2357(int v);
int next_smooth_2357(int n);
int next_even_smooth_2357(int n);
int previous_smooth_2357(int n);
int segment_length_from_seconds(float seconds, int max_samples, float delta);
int impulse_length(const Butterworth &f, int max_samples);
int filter_padding(const std::vector<Butterworth> &fs, int max_padding);
// |H|^2 at k = 0..nfft/2 (length nfft/2+1)
std::vector<float> power_response(const Butterworth &f, int nfft);

struct S2sLengths {
  int npts = 0, seg = 0, shift = 0, nstep = 0, nfft_out = 0, nspec = 0, nfft_filter = 0;
  float df_out = 0.f;
  int idx[4] = {0, 0, 0, 0};
};

// false with a logged reason on invalid input
bool derive_lengths(float sac_len, float delta, float win_len, float shift_len, float max_lag,
                    const std::vector<Butterworth> &filters, S2sLengths *out);

}  // namespace fxh
