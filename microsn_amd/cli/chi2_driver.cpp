// chi2_driver -- the reference's standalone multi-band chi2 driver
// (chi2/chi2_driver.cu:8-79, chi2_functions.cuh:93-190) on top of
// libmlc.so: constant fake data, a device run with the reference kernel's
// indexing (mu2 read at x, chi2_functions.cuh:52) and a host run of the CPU
// twin (mu2 read at y, chi2_functions.cuh:160-190), printed in the
// reference's order.  chi2_all is [N*y + x], mu is [n][f][t], f_obs and
// df_obs are [f][t].
//
// The likelihood sum exp(-chi2) over all N^2 pairs is accumulated in FP64;
// the reference accumulates it in float, which stops growing at 1024 for
// this data (the exact value is 4508.1).  Optional arguments N F T replace
// the reference's sizes 10000 7 10.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "../../include/mlc.h"

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// getChi2CPU (chi2_functions.cuh:160-190): chi2 of pair (x, y) compares mu1 of
// x with mu2 of y; one running float sum over filters and epochs.
double chi2_cpu(const std::vector<float>& f_obs, const std::vector<float>& df_obs,
                const std::vector<float>& mu1, const std::vector<float>& mu2, int N, int F, int T) {
  double like = 0.0;
  const int FT = F * T;
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y) {
      float c = 0.0f;
      for (int i = 0; i < FT; ++i) {
        const float r = (f_obs[i] - mu1[(size_t)x * FT + i] / mu2[(size_t)y * FT + i]) / df_obs[i];
        c += r * r / 2.0f;
      }
      like += std::exp(-(double)c);
    }
  return like;
}

}  // namespace

int main(int argc, char* argv[]) {
  int N = 10000, F = 7, T = 10;
  if (argc == 4) {
    N = std::atoi(argv[1]);
    F = std::atoi(argv[2]);
    T = std::atoi(argv[3]);
  } else if (argc != 1) {
    std::cerr << "usage: chi2_driver [N F T]" << std::endl;
    return 1;
  }
  if (N <= 0 || F <= 0 || T <= 0) {
    std::cerr << "Error: N, F and T must be positive" << std::endl;
    return 1;
  }
  const size_t FT = (size_t)F * T, NFT = (size_t)N * FT;
  std::vector<float> f_obs(FT, 0.97f), df_obs(FT, 1.1f), mu1(NFT, 0.42f), mu2(NFT, 1.1f);

  mlc_ctx* ctx = nullptr;
  if (mlc_ctx_create(0, nullptr, &ctx) != MLC_OK) {
    std::cerr << "Error: mlc_ctx_create failed" << std::endl;
    return 1;
  }
  float *df, *ddf, *dm1, *dm2, *dchi2;
  int rc = mlc_malloc(ctx, FT * sizeof(float), (void**)&df) | mlc_malloc(ctx, FT * sizeof(float), (void**)&ddf) |
           mlc_malloc(ctx, NFT * sizeof(float), (void**)&dm1) | mlc_malloc(ctx, NFT * sizeof(float), (void**)&dm2) |
           mlc_malloc(ctx, (size_t)N * N * sizeof(float), (void**)&dchi2);
  if (rc != MLC_OK) {
    std::cerr << "Error: " << mlc_last_error(ctx) << std::endl;
    mlc_ctx_destroy(ctx);
    return 1;
  }
  mlc_memcpy_h2d(ctx, df, f_obs.data(), FT * sizeof(float));
  mlc_memcpy_h2d(ctx, ddf, df_obs.data(), FT * sizeof(float));
  mlc_memcpy_h2d(ctx, dm1, mu1.data(), NFT * sizeof(float));
  mlc_memcpy_h2d(ctx, dm2, mu2.data(), NFT * sizeof(float));

  std::cout << "GPU run:" << std::endl;
  double like = 0.0;
  auto t0 = std::chrono::steady_clock::now();
  rc = mlc_chi2_multiband(ctx, df, ddf, dm1, dm2, N, F, T, 1, dchi2, &like);
  if (rc != MLC_OK) {
    std::cerr << "Error: " << mlc_last_error(ctx) << std::endl;
    mlc_ctx_destroy(ctx);
    return 1;
  }
  // the reference prints the time after the chi2 kernel and again after its
  // host summation; here the per-block sums of exp(-chi2) come out of the
  // chi2 launch itself, so both times end with the same call
  const double t_chi2 = seconds_since(t0);
  std::cout << "Elapsed time: " << t_chi2 << std::endl;
  std::cout << "Elapsed time: " << seconds_since(t0) << std::endl;
  std::cout << "Likelihood: " << like << std::endl;
  size_t free_b = 0, total_b = 0;
  mlc_mem_info(ctx, &free_b, &total_b);
  std::cout << "number of free bytes in gpu memory: " << free_b << "\n";
  for (void* p : {(void*)df, (void*)ddf, (void*)dm1, (void*)dm2, (void*)dchi2}) mlc_free(ctx, p);
  mlc_ctx_destroy(ctx);

  std::cout << "CPU run:" << std::endl;
  t0 = std::chrono::steady_clock::now();
  like = chi2_cpu(f_obs, df_obs, mu1, mu2, N, F, T);
  std::cout << "Elapsed time: " << seconds_since(t0) << std::endl;
  std::cout << "Likelihood: " << like << std::endl;
  return 0;
}
