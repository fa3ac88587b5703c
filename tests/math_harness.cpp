// math_harness.cpp -- CPU driver for the scalar helpers of
// microsn_amd/csrc/mlc_math.hpp (the header the gfx950 kernels include).
// Built by tests/test_fft_core.py with g++.  Test infrastructure only.
#include "../microsn_amd/csrc/mlc_math.hpp"

extern "C" void h_ln(const double* x, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = mlc::ln_pos(x[i]);
}

extern "C" void h_ln_f32(const float* x, int n, double* out) {
  mlc::LnEntry tab[mlc::kLnTab];
  for (int i = 0; i < mlc::kLnTab; ++i) tab[i] = mlc::ln_entry(i);
  for (int i = 0; i < n; ++i) out[i] = mlc::ln_f32(x[i], tab);
}
