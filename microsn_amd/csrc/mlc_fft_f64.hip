// mlc_fft_f64.hip -- FP64 instantiation of the FFT and convolution kernels
// (mlc_fft_impl.hpp), plus the profile evaluation used by mlc_profile.
#include "mlc_fft_impl.hpp"

namespace mlc {
MLC_FFT_INSTANTIATE(double)

void fft_profile(mlc_ctx* ctx, const ProfileParams& P, double* prof, double* dprof) {
  MLC_LAUNCH(ctx, "profile", k_profile, dim3(1), dim3(1024), 0, P, prof, dprof);
}
}  // namespace mlc
