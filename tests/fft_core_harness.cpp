// fft_core_harness.cpp -- CPU driver for the butterflies, Stockham index map
// and real-FFT pre/post steps of microsn_amd/csrc/mlc_fft_core.hpp (the same
// header the gfx950 kernels include).  Built by tests/test_fft_core.py with
// g++ and compared against numpy's FFT.  Test infrastructure only.
#include <cmath>
#include <type_traits>
#include <vector>

#include "../microsn_amd/csrc/mlc_ctfft.hpp"
#include "../microsn_amd/csrc/mlc_fft_core.hpp"

using namespace mlc;

// Runs all stages ping-ponging between two buffers.  With `quarter` the
// twiddles come from a quarter-wave table, as in the gfx950 kernels.
template <class T, bool INV>
static void run(std::vector<cpx<T>>& a, int L, bool quarter = false) {
  Plan p;
  if (!make_plan(L, &p)) return;
  std::vector<cpx<T>> tw(L);
  for (int m = 0; m < L; ++m) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * m / L;
    tw[m] = mk<T>((T)std::cos(ang), (T)std::sin(ang));
  }
  std::vector<cpx<T>> x(a), y(L);
  TableTw<T> full{tw.data()};
  QuarterTw<T> qtw{tw.data(), L / 4};
  int Ns = 1;
  for (int s = 0; s < p.n; ++s) {
    const int R = p.r[s];
    for (int j = 0; j < L / R; ++j) {
      cpx<T> v[16];
#define H_CASE(RR)                                                                      \
  case RR:                                                                              \
    if (quarter && L % 4 == 0) stage_load<T, RR, INV>(x.data(), L, Ns, j, qtw, v);     \
    else stage_load<T, RR, INV>(x.data(), L, Ns, j, full, v);                          \
    stage_store<T, RR>(y.data(), Ns, j, v);                                            \
    break;
      switch (R) {
        H_CASE(2) H_CASE(3) H_CASE(4) H_CASE(5) H_CASE(8) H_CASE(16)
      }
#undef H_CASE
    }
    x.swap(y);
    Ns *= R;
  }
  a = x;
}

extern "C" {

// In-place complex FFT of n interleaved doubles; returns -1 for an
// unsupported length.
int h_fft(double* data, int L, int inverse, int quarter) {
  Plan p;
  if (!make_plan(L, &p)) return -1;
  std::vector<cpx<double>> a(L);
  for (int i = 0; i < L; ++i) a[i] = mk(data[2 * i], data[2 * i + 1]);
  if (inverse) run<double, true>(a, L, quarter != 0);
  else run<double, false>(a, L, quarter != 0);
  for (int i = 0; i < L; ++i) {
    data[2 * i] = a[i].x;
    data[2 * i + 1] = a[i].y;
  }
  return 0;
}

// Real forward FFT of N reals -> N/2+1 complex (interleaved), through the
// half-length complex transform and r2c_post.
int h_rfft(const double* x, int N, double* X) {
  int M = N / 2;
  std::vector<cpx<double>> z(M);
  for (int n = 0; n < M; ++n) z[n] = mk(x[2 * n], x[2 * n + 1]);
  Plan p;
  if (N % 2 || !make_plan(M, &p)) return -1;
  run<double, false>(z, M);
  for (int k = 0; k <= M / 2; ++k) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / N;
    cpx<double> w = mk((double)std::cos(ang), (double)std::sin(ang));
    cpx<double> Xk, Xm;
    r2c_post(z[k % M], z[(M - k) % M], w, &Xk, &Xm);
    X[2 * k] = Xk.x;
    X[2 * k + 1] = Xk.y;
    X[2 * (M - k)] = Xm.x;
    X[2 * (M - k) + 1] = Xm.y;
  }
  return 0;
}

// Unnormalised inverse real FFT: N/2+1 complex -> N reals.
int h_irfft(const double* X, int N, double* x) {
  int M = N / 2;
  Plan p;
  if (N % 2 || !make_plan(M, &p)) return -1;
  std::vector<cpx<double>> z(M);
  for (int k = 0; k <= M / 2; ++k) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / N;
    cpx<double> w = mk((double)std::cos(ang), (double)std::sin(ang));
    cpx<double> Zk, Zm;
    c2r_pre(mk(X[2 * k], X[2 * k + 1]), mk(X[2 * (M - k)], X[2 * (M - k) + 1]), w, &Zk, &Zm);
    z[k] = Zk;
    if (k > 0) z[M - k] = Zm;
  }
  run<double, true>(z, M);
  for (int n = 0; n < M; ++n) {
    x[2 * n] = z[n].x;
    x[2 * n + 1] = z[n].y;
  }
  return 0;
}
}

// ---- compile-time Stockham (mlc_ctfft.hpp), every thread of a stage run in
// turn before the barrier and again after it
template <class T, int N, int NS, bool INV>
static void ct_host_stages(cpx<T>* buf, const cpx<T>* lo, const cpx<T>* hi) {
  if constexpr (NS < N) {
    using S = CtStage<T, N, NS, INV>;
    constexpr int G = CtGeom<N>::G;
    std::vector<typename std::remove_reference<typename S::Regs>::type> regs(G);
    for (int t = 0; t < G; ++t) S::load(buf, lo, hi, t, regs[t]);
    for (int t = 0; t < G; ++t) S::store(buf, t, regs[t]);
    ct_host_stages<T, N, NS * S::R, INV>(buf, lo, hi);
  }
}

template <int N>
static void ct_host(double* data, int inverse) {
  using Gm = CtGeom<N>;
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double sg = inverse ? 1.0L : -1.0L;
  // one table as the kernels hold it: lo | hi | t2 (the second stage's)
  std::vector<cpx<double>> tw(Gm::NTW), buf(Gm::LDS);
  cpx<double>* lo = tw.data();
  cpx<double>* hi = lo + Gm::NLO;
  cpx<double>* t2 = hi + Gm::NHI;
  for (int l = 0; l < Gm::NLO; ++l)
    lo[l] = mk((double)std::cos(sg * 2 * pi * l / N), (double)std::sin(sg * 2 * pi * l / N));
  for (int h = 0; h < Gm::NHI; ++h) {
    long double a = sg * 2 * pi * (long double)h * Gm::NLO / N;
    hi[h] = mk((double)std::cos(a), (double)std::sin(a));
  }
  for (int r = 0; r < 16; ++r)
    for (int k = 0; k < 16; ++k) {
      long double a = sg * 2 * pi * (long double)(k * r) / 256;
      t2[16 * r + k] = mk((double)std::cos(a), (double)std::sin(a));
    }
  for (int i = 0; i < N; ++i) buf[pad16(i)] = mk(data[2 * i], data[2 * i + 1]);
  if (inverse) ct_host_stages<double, N, 1, true>(buf.data(), lo, hi);
  else ct_host_stages<double, N, 1, false>(buf.data(), lo, hi);
  for (int i = 0; i < N; ++i) {
    data[2 * i] = buf[pad16(i)].x;
    data[2 * i + 1] = buf[pad16(i)].y;
  }
}

extern "C" int h_ctfft(double* data, int N, int inverse) {
  switch (N) {
    case 256: ct_host<256>(data, inverse); return 0;
    case 512: ct_host<512>(data, inverse); return 0;
    case 1024: ct_host<1024>(data, inverse); return 0;
    case 2048: ct_host<2048>(data, inverse); return 0;
    case 4096: ct_host<4096>(data, inverse); return 0;
    case 8192: ct_host<8192>(data, inverse); return 0;
    case 16384: ct_host<16384>(data, inverse); return 0;
  }
  return -1;
}
