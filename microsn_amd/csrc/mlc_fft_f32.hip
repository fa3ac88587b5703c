// mlc_fft_f32.hip -- FP32 instantiation of the FFT and convolution kernels
// (mlc_fft_impl.hpp), plus the size checks shared by both precisions.
#include "mlc_fft_impl.hpp"

namespace mlc {
MLC_FFT_INSTANTIATE(float)

void fft_check_dims(int Nx, int Ny) {
  MLC_REQUIRE(Nx >= 2 && Ny >= 1, "map size must be at least 2 x 1");
  MLC_REQUIRE(Nx % 2 == 0, "map width Nx must be even");
  MLC_REQUIRE((int64_t)Ny * (Nx / 2 + 1) < (int64_t)1 << 30, "map too large for 32-bit offsets");
  (void)plan_for(Nx / 2, "row FFT");
  (void)plan_for(Ny, "column FFT");
}


}  // namespace mlc
