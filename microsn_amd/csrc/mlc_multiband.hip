// mlc_multiband.hip -- the chi2_driver path: kernelChi2 + getChi2GPU /
// getChi2CPU (chi2/chi2_functions.cuh:35-190).
//
//   chi2_all[N*y+x] = sum_f sum_t ((f_obs[f,t] - mu1[x,f,t]/mu2[m,f,t]) / df[f,t])^2 / 2
//
// FP32, built with -ffp-contract=off and IEEE division so that every pair is
// bit-identical to the reference's float arithmetic in the chosen order:
//   mode 0 (getChi2CPU, 160-190):  m = y, one running float over (f, t)
//   mode 1 (kernelChi2, 35-68):    m = x, a float per filter over t, the
//                                  filter sums added in filter order (the
//                                  reference adds them with atomics in an
//                                  unspecified order)
// The likelihood sum_{x,y} exp(-chi2_all) is reduced in FP64 (the reference
// accumulates it in one float on the host, which stagnates; see the
// chi2_driver output of the host code).
//
// Tiling: a workgroup owns 64 x (one per lane) by 16 y (4 per thread) and
// stages the observation rows it needs in LDS in chunks of 64 (f,t) terms.
#include <vector>

#include "mlc_internal.hpp"

using namespace mlc;

namespace {

constexpr int kTX = 64, kTY = 16, kQ = 64;

template <int MODE>
__global__ __launch_bounds__(256) void k_multiband(const float* __restrict__ fobs,
                                                   const float* __restrict__ dfobs,
                                                   const float* __restrict__ mu1,
                                                   const float* __restrict__ mu2, int N, int F,
                                                   int T, float* __restrict__ chi2_all,
                                                   double* __restrict__ part) {
  __shared__ float s1[kTX][kQ + 1];
  __shared__ float s2[MODE == 0 ? kTY : kTX][kQ + 1];
  __shared__ float sf[kQ], sdf[kQ];
  __shared__ double red[256];
  const int t = threadIdx.x, lane = t & 63, ty = t >> 6;
  const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY;
  const int x = x0 + lane;
  const int FT = F * T;
  float chi[4] = {0.f, 0.f, 0.f, 0.f}, prt[4] = {0.f, 0.f, 0.f, 0.f};
  for (int q0 = 0; q0 < FT; q0 += kQ) {
    const int nq = min(kQ, FT - q0);
    __syncthreads();
    for (int i = t; i < kTX * kQ; i += 256) {
      int r = i / kQ, q = i % kQ, xx = x0 + r;
      s1[r][q] = (xx < N && q < nq) ? mu1[(size_t)xx * FT + q0 + q] : 1.f;
      if (MODE == 1) s2[r][q] = (xx < N && q < nq) ? mu2[(size_t)xx * FT + q0 + q] : 1.f;
    }
    if (MODE == 0)
      for (int i = t; i < kTY * kQ; i += 256) {
        int r = i / kQ, q = i % kQ, yy = y0 + r;
        s2[r][q] = (yy < N && q < nq) ? mu2[(size_t)yy * FT + q0 + q] : 1.f;
      }
    if (t < kQ) {
      sf[t] = t < nq ? fobs[q0 + t] : 0.f;
      sdf[t] = t < nq ? dfobs[q0 + t] : 1.f;
    }
    __syncthreads();
    for (int q = 0; q < nq; ++q) {
      const float a = s1[lane][q], fo = sf[q], df = sdf[q];
      const bool last_t = ((q0 + q) % T) == T - 1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float b = (MODE == 0) ? s2[ty + 4 * j][q] : s2[lane][q];
        float tmp = (fo - a / b) / df;
        if (MODE == 0) {
          chi[j] += tmp * tmp / 2;
        } else {
          prt[j] += tmp * tmp / 2;
          if (last_t) {
            chi[j] += prt[j];
            prt[j] = 0.f;
          }
        }
      }
    }
  }
  double ls = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int y = y0 + ty + 4 * j;
    if (x < N && y < N) {
      chi2_all[(size_t)N * y + x] = chi[j];
      ls += exp(-(double)chi[j]);
    }
  }
  red[t] = ls;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void k_sum_ordered(const double* __restrict__ part, int64_t n,
                                                      double* __restrict__ out) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  // contiguous slice per thread, then a fixed tree: deterministic
  const int64_t per = (n + 1023) / 1024;
  double s = 0.0;
  for (int64_t i = t * per; i < min<int64_t>(n, (t + 1) * per); ++i) s += part[i];
  red[t] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[0] = red[0];
}

}  // namespace

extern "C" int mlc_chi2_multiband(mlc_ctx* ctx, const float* d_f_obs, const float* d_df_obs,
                                  const float* d_mu1, const float* d_mu2, int N, int F, int T,
                                  int mode, float* d_chi2_all, double* likelihood) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(N > 0 && F > 0 && T > 0, "mlc_chi2_multiband: bad sizes");
    MLC_REQUIRE(mode == 0 || mode == 1, "mlc_chi2_multiband: mode must be 0 or 1");
    MLC_REQUIRE(d_f_obs && d_df_obs && d_mu1 && d_mu2 && d_chi2_all && likelihood,
                "mlc_chi2_multiband: null pointer");
    MLC_REQUIRE((int64_t)N * N < ((int64_t)1 << 40), "mlc_chi2_multiband: N too large");
    dim3 grid(cdiv(N, kTX), cdiv(N, kTY));
    const int64_t nparts = (int64_t)grid.x * grid.y;
    Workspace ws(ctx);
    int ip = ws.add(sizeof(double) * nparts), io = ws.add(sizeof(double));
    ws.commit();
    if (mode == 0)
      MLC_LAUNCH(ctx, "multiband_chi2", k_multiband<0>, grid, dim3(256), 0, d_f_obs, d_df_obs, d_mu1,
                 d_mu2, N, F, T, d_chi2_all, ws.get<double>(ip));
    else
      MLC_LAUNCH(ctx, "multiband_chi2", k_multiband<1>, grid, dim3(256), 0, d_f_obs, d_df_obs, d_mu1,
                 d_mu2, N, F, T, d_chi2_all, ws.get<double>(ip));
    MLC_LAUNCH(ctx, "multiband_likelihood", k_sum_ordered, dim3(1), dim3(1024), 0,
               ws.get<double>(ip), nparts, ws.get<double>(io));
    d2h(ctx, likelihood, ws.get<double>(io), sizeof(double));
    sync(ctx);
  });
}
