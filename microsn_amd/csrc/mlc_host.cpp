// mlc_host.cpp -- host-side steps of the reference driver that feed the
// device path: sample locations, the chi2 set-up tables, MPD bin edges and
// the trapezium integral.  Plain C++ (no device code), exported through the
// same C ABI as the kernels.
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/mlc_host.h"

extern "C" {

// locations_on_grid (locations.cpp:6-14).  The spacing is an integer
// division, as in the reference, then used as a double.
int mlc_locations_on_grid(int Nside, int Nmap, int Noff, int* x, int* y) {
  if (Nside <= 0 || Noff < 0 || Nmap - 2 * Noff <= 0 || !x || !y) return 1;
  const double step = (double)((Nmap - 2 * Noff) / Nside);
  for (int row = 0; row < Nside; ++row) {
    const int iy = Noff + (int)std::floor(step / 2.0 + row * step);
    for (int col = 0; col < Nside; ++col) {
      x[row * Nside + col] = Noff + (int)std::floor(step / 2.0 + col * step);
      y[row * Nside + col] = iy;
    }
  }
  return 0;
}

// locations_random (locations.cpp:17-25): mt19937(seed) and an inclusive
// uniform integer in [0, Nmap - 2 Noff], x drawn before y for each location.
// The integer distribution is libstdc++'s (the reference's toolchain).
int mlc_locations_random(int Nloc, int Nmap, int Noff, int seed, int* x, int* y) {
  if (Nloc < 0 || (Nloc > 0 && (!x || !y)) || Noff < 0 || Nmap - 2 * Noff < 0) return 1;
  std::mt19937 gen(seed);
  std::uniform_int_distribution<unsigned> pick(0, (unsigned)(Nmap - 2 * Noff));
  for (int i = 0; i < Nloc; ++i) {
    x[i] = Noff + (int)pick(gen);
    y[i] = Noff + (int)pick(gen);
  }
  return 0;
}

// find_indices (cpu_functions.cpp:28-46): the left index j of the size
// interval x[j] <= t < x[j+1] holding each point, -1 beyond the last size.
// For ascending t inside the grid this is the reference's forward walk; the
// cases it leaves unset or reads past t[Nd-1] for are defined here: the last
// interval is closed (t == x[Nprof-1] gives Nprof-2) and t < x[0] gives -1.
int mlc_find_indices(const double* t, const double* x, int Nd, int Nprof, int* ind) {
  if (Nd < 0 || Nprof < 1 || (Nd > 0 && (!t || !x || !ind))) return 1;
  for (int i = 0; i < Nd; ++i) {
    const double v = t[i];
    ind[i] = -1;
    if (Nprof < 2 || !(v >= x[0]) || v > x[Nprof - 1]) continue;
    int j = 0;
    while (j < Nprof - 2 && !(v < x[j + 1])) ++j;
    ind[i] = j;
  }
  return 0;
}

// setup_chi2_calculation (cpu_functions.cpp:80-144).  Rfac = 23.92865 sqrt(DS DLS/DL)
// enters through Dfac = sqrt(DS DLS / DL); t2x = V 8.64/(Rfac sqrt(M)).
int mlc_setup_chi2(double MA, double MB, double V, double Dfac, int Nd, const double* d,
                   const double* t, const double* s, int Nprof, const double* x,
                   int quirk_const_term, int* indA, int* indB, double* facA, double* facB,
                   double* new_d, double* new_s, int* Njp, double* const_term) {
  if (Nd < 0 || Nprof < 1 || !Njp || !const_term) return 1;
  if (Nd > 0 && (!d || !t || !s || !x || !indA || !indB || !facA || !facB || !new_d || !new_s))
    return 1;
  const double Rfac = 23.92865 * Dfac;
  const double kA = V * 8.64 / (Rfac * std::sqrt(MA));
  const double kB = V * 8.64 / (Rfac * std::sqrt(MB));
  std::vector<double> xa(Nd), xb(Nd);
  for (int i = 0; i < Nd; ++i) {
    xa[i] = kA * t[i];
    xb[i] = kB * t[i];
  }
  std::vector<int> ia(Nd), ib(Nd);
  mlc_find_indices(xa.data(), x, Nd, Nprof, ia.data());
  mlc_find_indices(xb.data(), x, Nd, Nprof, ib.data());
  int n = Nd;  // first point where both images have left the size grid
  for (int i = 0; i < Nd; ++i)
    if (ia[i] == -1 && ib[i] == -1) {
      n = i;
      break;
    }
  for (int i = 0; i < n; ++i) {
    indA[i] = ia[i];
    indB[i] = ib[i];
    new_d[i] = d[i];
    new_s[i] = s[i];
    facA[i] = ia[i] == -1 ? 0.0 : (xa[i] - x[ia[i]]) / (x[ia[i] + 1] - x[ia[i]]);
    facB[i] = ib[i] == -1 ? 0.0 : (xb[i] - x[ib[i]]) / (x[ib[i] + 1] - x[ib[i]]);
  }
  // constant term of the points left out; the reference starts its loop at
  // Nd - Njp (cpu_functions.cpp:138), reproduced when quirk_const_term != 0
  double c = 0.0;
  if (n != Nd)
    for (int i = quirk_const_term ? Nd - n : n; i < Nd; ++i) c += std::pow((d[i] - 1) / s[i], 2);
  *Njp = n;
  *const_term = c;
  return 0;
}

// Mpd::Mpd(Nbins, min, max, type) bin right edges (mpd.cpp:4-21).
int mlc_mpd_bins(int Nbins, double min, double max, int log_scale, double* bins) {
  if (Nbins <= 0 || !bins || !(max > min)) return 1;
  if (log_scale) {
    if (!(min > 0) || !(max > 0)) return 1;
    const double lo = std::log10(min), step = (std::log10(max) - lo) / Nbins;
    for (int i = 0; i < Nbins; ++i) bins[i] = std::pow(10, lo + (i + 1) * step);
  } else {
    const double step = (max - min) / Nbins;
    for (int i = 0; i < Nbins; ++i) bins[i] = min + (i + 1) * step;
  }
  return 0;
}

// Mpd::normalize (mpd.cpp:63-68 with integrate 47-61): counts /= trapezoid area.
int mlc_mpd_normalize(int N, const double* bins, double* counts) {
  if (N <= 0 || !bins || !counts) return 1;
  double area = 0.0;
  for (int i = 1; i < N; ++i) area += (counts[i] + counts[i - 1]) * (bins[i] - bins[i - 1]) / 2.0;
  for (int i = 0; i < N; ++i) counts[i] /= area;
  return 0;
}

// Mpd::divide_by (mpd.cpp:71-107): PDF of the ratio z = mu_A/mu_B,
// P(z) = int mu_B P_A(z mu_B) P_B(mu_B) dmu_B with P_A linearly interpolated
// and 1e-8 outside its range, trapezoid rule over the B bins.
int mlc_mpd_divide_by(int NA, const double* binsA, const double* countsA, int NB,
                      const double* binsB, const double* countsB, int NR, const double* binsR,
                      double* countsR) {
  if (NA < 2 || NB < 1 || NR < 1 || !binsA || !countsA || !binsB || !countsB || !binsR || !countsR)
    return 1;
  std::vector<double> f(NB);
  for (int r = 0; r < NR; ++r) {
    const double z = binsR[r];
    for (int b = 0; b < NB; ++b) {
      const double mu = binsB[b], arg = z * mu;
      double pa = 1e-8;
      if (arg > binsA[0] && arg < binsA[NA - 1]) {
        int k = 1;
        while (!(binsA[k] > arg)) ++k;
        pa = countsA[k - 1] + (countsA[k] - countsA[k - 1]) * (arg - binsA[k - 1]) / (binsA[k] - binsA[k - 1]);
      }
      f[b] = mu * pa * countsB[b];
    }
    double sum = 0.0;
    for (int b = 1; b < NB; ++b) sum += (f[b - 1] + f[b]) * (binsB[b] - binsB[b - 1]) / 2.0;
    countsR[r] = sum;
  }
  return 0;
}

// trapezium (integration.cpp:3-9): integral of f1*f2 over x.
double mlc_trapezium(int N, const double* f1, const double* f2, const double* x) {
  double acc = 0.0;
  for (int i = 1; i < N; ++i) acc += (f1[i - 1] * f2[i - 1] + f1[i] * f2[i]) * (x[i] - x[i - 1]) / 2.0;
  return acc;
}

}  // extern "C"
