// mlc_ctfft.hpp -- Stockham FFT of a compile-time power-of-two length held
// in LDS, for the hot row transforms.  All index arithmetic folds to shifts
// and masks; N/16 threads each own 16 elements per stage: radix-16 stages,
// then one radix-8/4/2 stage for the remaining factor.
//
// Stage after Ns = product of the earlier radices, radix R, butterfly j:
//   inputs   x[j + r N/R], twiddled by w^(r k N/(Ns R)), k = j mod Ns
//   outputs  y[(j/Ns) Ns R + k + q Ns]
// (the map of mlc_fft_core.hpp).  Each stage reads all of its inputs into
// registers, synchronises, then writes, so one buffer serves in place.
// Twiddles w^m = exp(-+2 pi i m/N) come from two short tables, w^m =
// hi[m >> LB] * lo[m & (2^LB - 1)], one complex product per twiddle, except
// in the second stage (Ns = 16), whose 256 twiddles w^(k r N/256), k, r <
// 16, are a table of their own stored right after hi (t2[16 r + k], so the
// 16 values of k read together lie in consecutive banks).
// LDS indices are padded by one element per 16 (pad16) so the strided
// writes of the first stages do not pile onto a few banks.
#pragma once

#include "mlc_fft_core.hpp"

namespace mlc {

constexpr int ct_log2(int n) { return n <= 1 ? 0 : 1 + ct_log2(n >> 1); }

template <int N>
struct CtGeom {
  static_assert(N >= 256 && (N & (N - 1)) == 0, "power-of-two length >= 256");
  static constexpr int G = N / 16;           // threads per transform
  static constexpr int LB = ct_log2(N) / 2;  // bits of the low twiddle table
  static constexpr int NLO = 1 << LB, NHI = N >> LB;
  static constexpr int NT2 = 256, NTW = NLO + NHI + NT2;  // lo | hi | t2
  static constexpr int LDS = N + N / 16;     // padded buffer length (elements)
};

MLC_HD int pad16(int i) { return i + (i >> 4); }

template <class T, int R, bool INV>
MLC_HD void ct_dft(cpx<T>* v) {
  if constexpr (R == 16) dft16<T, INV>(v);
  else if constexpr (R == 8) dft8<T, INV>(v);
  else if constexpr (R == 4) dft4<T, INV>(v);
  else dft2<T, INV>(v);
}

// One stage, split into the part before the barrier (load, twiddle,
// butterfly into registers) and the part after it (store), so that a CPU
// harness can run the same code thread by thread.
template <class T, int N, int NS, bool INV>
struct CtStage {
  using Gm = CtGeom<N>;
  static constexpr int REM = N / NS;
  static constexpr int R = REM >= 16 ? 16 : REM;
  static constexpr int NB = N / R;    // butterflies
  static constexpr int BPT = 16 / R;  // butterflies per thread
  static constexpr int LNS = ct_log2(NS);
  using Regs = cpx<T>[BPT][R];

  // pad16(j + r NB) = pad16(j) + r (NB + NB/16) and pad16(base + q NS) =
  // pad16(base) + q (NS + NS/16) when 16 divides NB and NS (NS = 1: base is
  // a multiple of 16 and q < 16), so the padded offsets are constants the
  // LDS instructions carry and only pad16(j), pad16(base) are computed.
  static_assert(NB % 16 == 0 && (NS == 1 || NS % 16 == 0), "padded strides are constants");
  static constexpr int PNB = NB + NB / 16, PNS = NS == 1 ? 1 : NS + NS / 16;

  MLC_HD static void load(const cpx<T>* buf, const cpx<T>* lo, const cpx<T>* hi, int t, Regs& v) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * Gm::G;
      const cpx<T>* src = buf + pad16(j);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = src[r * PNB];
      if constexpr (NS == 16) {
        const cpx<T>* t2 = hi + Gm::NHI + (j & 15);
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = mul(v[b][r], t2[16 * r]);
      } else if constexpr (NS > 1) {
        const int k = j & (NS - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int m = k * r * (N / (NS * R));
          const cpx<T> w = mul(hi[m >> Gm::LB], lo[m & (Gm::NLO - 1)]);
          v[b][r] = mul(v[b][r], w);
        }
      }
      ct_dft<T, R, INV>(v[b]);
    }
  }

  MLC_HD static void store(cpx<T>* buf, int t, const Regs& v) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * Gm::G;
      const int base = ((j >> LNS) << (LNS + ct_log2(R))) + (j & (NS - 1));
      cpx<T>* dst = buf + pad16(base);
#pragma unroll
      for (int q = 0; q < R; ++q) dst[q * PNS] = v[b][q];
    }
  }
};

#ifdef __HIPCC__
template <class T, int N, int NS, bool INV>
__device__ __forceinline__ void ct_stage(cpx<T>* buf, const cpx<T>* lo, const cpx<T>* hi, int t) {
  using S = CtStage<T, N, NS, INV>;
  typename S::Regs v;
  S::load(buf, lo, hi, t, v);
  __syncthreads();
  S::store(buf, t, v);
  __syncthreads();
}

template <class T, int N, int NS, bool INV>
__device__ __forceinline__ void ct_stages(cpx<T>* buf, const cpx<T>* lo, const cpx<T>* hi, int t) {
  if constexpr (NS < N) {
    ct_stage<T, N, NS, INV>(buf, lo, hi, t);
    constexpr int R = (N / NS) >= 16 ? 16 : (N / NS);
    ct_stages<T, N, NS * R, INV>(buf, lo, hi, t);
  }
}

// In-place length-N transform of buf (padded indices, pad16), called by all
// CtGeom<N>::G threads of the group (t = 0..G-1) after buf is complete
// (the caller synchronises before).  Result in natural order at pad16(i).
template <class T, int N, bool INV>
__device__ __forceinline__ void ct_fft(cpx<T>* buf, const cpx<T>* lo, const cpx<T>* hi, int t) {
  ct_stages<T, N, 1, INV>(buf, lo, hi, t);
}
#endif  // __HIPCC__

}  // namespace mlc
