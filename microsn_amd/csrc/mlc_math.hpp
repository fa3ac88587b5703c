// mlc_math.hpp -- scalar FP64 helpers shared by host and device code.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MLC_MHD __host__ __device__ __forceinline__
#else
#define MLC_MHD inline
#endif

namespace mlc {

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi

// Natural logarithm of a positive, finite double (normal or subnormal) to
// within about two ulp; other inputs give unspecified values, so callers
// reject them themselves.  Classic argument reduction: x = 2^e m with
// m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m - 1)/(m + 1),
// |s| < 0.1716, summed through s^19 (the first omitted term is below
// 1e-17), and e ln 2 added in two parts (the high part has trailing zero
// bits, so e * kLn2Hi is exact for |e| < 2^11).  About 30 instructions on
// gfx950 against roughly 100 for the library log10.
MLC_MHD double ln_pos(double x) {
  int e;
  double m = std::frexp(x, &e);  // m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);  // m - 1 is exact
  const double z = s * s;
  double p = 1.0 / 19.0;
  p = std::fma(p, z, 1.0 / 17.0);
  p = std::fma(p, z, 1.0 / 15.0);
  p = std::fma(p, z, 1.0 / 13.0);
  p = std::fma(p, z, 1.0 / 11.0);
  p = std::fma(p, z, 1.0 / 9.0);
  p = std::fma(p, z, 1.0 / 7.0);
  p = std::fma(p, z, 1.0 / 5.0);
  p = std::fma(p, z, 1.0 / 3.0);
  const double s2 = 2.0 * s;               // exact
  const double lnm = std::fma(s2 * z, p, s2);
  const double de = (double)e;
  return std::fma(de, kLn2Hi, std::fma(de, kLn2Lo, lnm));
}

// Entry i of the ln_f32 table: C = 2 / (1 + (i + 1/2)/256) and T = -ln C.
struct alignas(16) LnEntry {
  double c, t;
};
constexpr int kLnTab = 256;

MLC_MHD LnEntry ln_entry(int i) {
  const double c = 2.0 / (1.0 + (i + 0.5) / kLnTab);
  return LnEntry{c, -ln_pos(c)};
}

// Natural logarithm, in FP64, of a positive finite float (normal or
// subnormal); other inputs give unspecified values.  Table reduction:
// x = f 2^e with f in [1/2, 1) (frexp), i = the top 8 fraction bits of f,
// r = f C_i - 1 with one rounding (|r| < 2^-9), and
//   ln x = e ln 2 + T_i + ln(1 + r),
// ln(1 + r) through r^5 (the first omitted term is below 2^-54 / 6).  The
// error is about one ulp of the result plus 2e-16 absolute (the table
// entry's own).  About a third of the instructions of ln_pos on gfx950; the
// caller holds the 256 entries of ln_entry, in LDS on the device.
MLC_MHD double ln_f32(float x, const LnEntry* tab) {
#ifdef __HIP_DEVICE_COMPILE__
  const float f = __builtin_amdgcn_frexp_mantf(x);
  const int e = __builtin_amdgcn_frexp_expf(x);
#else
  int e;
  const float f = std::frexp(x, &e);
#endif
  const LnEntry E = tab[(__builtin_bit_cast(unsigned, f) >> 15) & (kLnTab - 1)];
  const double r = std::fma((double)f, E.c, -1.0);
  double a = std::fma(r, 0.2, -0.25);
  a = std::fma(r, a, 1.0 / 3.0);
  a = std::fma(r, a, -0.5);
  const double p = std::fma(r * r, a, r);
  const double de = (double)e;
  return std::fma(de, kLn2Hi, E.t) + std::fma(de, kLn2Lo, p);
}

}  // namespace mlc
