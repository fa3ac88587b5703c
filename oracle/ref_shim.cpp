// ref_shim.cpp -- extern "C" entry points onto the reference's own CPU
// sources (locations.cpp, mpd.cpp, integration.cpp), compiled unchanged from
// the reference tree by oracle/Makefile into oracle/_ref/libmlcref.so.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement.
#include <cstring>
#include "locations.hpp"
#include "mpd.hpp"
#include "integration.hpp"

extern "C" {
void ref_locations_on_grid(int Nside, int Nmap, int Noff, int* x, int* y) {
  locations_on_grid(Nside, Nmap, Noff, x, y);
}
void ref_locations_random(int Nloc, int Nmap, int Noff, int seed, int* x, int* y) {
  locations_random(Nloc, Nmap, Noff, seed, x, y);
}
void ref_mpd_bins(int Nbins, double min, double max, int log_scale, double* bins) {
  Mpd m(Nbins, min, max, log_scale ? "log" : "linear");
  std::memcpy(bins, m.bins, sizeof(double) * Nbins);
}
void ref_mpd_normalize(int N, const double* bins, double* counts) {
  Mpd m(N);
  std::memcpy(m.bins, bins, sizeof(double) * N);
  std::memcpy(m.counts, counts, sizeof(double) * N);
  m.normalize();
  std::memcpy(counts, m.counts, sizeof(double) * N);
}
void ref_mpd_divide_by(int NA, const double* binsA, const double* countsA, int NB,
                       const double* binsB, const double* countsB, int NR,
                       const double* binsR, double* countsR) {
  Mpd A(NA), B(NB), R(NR);
  std::memcpy(A.bins, binsA, sizeof(double) * NA);
  std::memcpy(A.counts, countsA, sizeof(double) * NA);
  std::memcpy(B.bins, binsB, sizeof(double) * NB);
  std::memcpy(B.counts, countsB, sizeof(double) * NB);
  std::memcpy(R.bins, binsR, sizeof(double) * NR);
  A.divide_by(&B, &R);
  std::memcpy(countsR, R.counts, sizeof(double) * NR);
}
double ref_trapezium(int N, double* f1, double* f2, double* x) { return trapezium(N, f1, f2, x); }
double ref_simpson(int N, double* f1, double* f2, double* x) { return simpson(N, f1, f2, x); }
}
