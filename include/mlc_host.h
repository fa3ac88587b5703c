/*
 * mlc_host.h -- host-only entry points of libmlc.so: the driver-side steps of
 * simulating_lcurves that prepare the device path's inputs.  Same C ABI
 * conventions as mlc.h (int status, 0 = success; host arrays only).
 * Citations are relative to the reference tree gvernard/microSN.
 */
#ifndef MLC_HOST_H
#define MLC_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* locations_on_grid (simulating_lcurves/src/locations.cpp:6-14;
 * include/locations.hpp): Nside*Nside pixel indices on a regular grid with
 * margin Noff; the spacing (Nmap-2*Noff)/Nside is an integer division. */
int mlc_locations_on_grid(int Nside, int Nmap, int Noff, int* x, int* y);

/* locations_random (locations.cpp:17-25): std::mt19937(seed) and
 * std::uniform_int_distribution<unsigned>(0, Nmap-2*Noff), x then y. */
int mlc_locations_random(int Nloc, int Nmap, int Noff, int seed, int* x, int* y);

/* find_indices (cpu_functions.cpp:28-46; cpu_functions.hpp:9). */
int mlc_find_indices(const double* t, const double* x, int Nd, int Nprof, int* ind);

/* setup_chi2_calculation (cpu_functions.cpp:80-144; cpu_functions.hpp:10)
 * for masses MA, MB and velocity V (the reference's M[0], M[1], V[0]).
 * Outputs (capacity Nd): indA, indB, facA, facB, new_d, new_s; *Njp and the
 * constant term.  quirk_const_term != 0 starts the constant-term loop at
 * Nd-Njp as the reference does (cpu_functions.cpp:138); 0 starts at Njp. */
int mlc_setup_chi2(double MA, double MB, double V, double Dfac, int Nd, const double* d,
                   const double* t, const double* s, int Nprof, const double* x,
                   int quirk_const_term, int* indA, int* indB, double* facA, double* facB,
                   double* new_d, double* new_s, int* Njp, double* const_term);

/* Mpd(Nbins, min, max, "log"|"linear") right bin edges (mpd.cpp:4-21). */
int mlc_mpd_bins(int Nbins, double min, double max, int log_scale, double* bins);
/* Mpd::normalize (mpd.cpp:47-68). */
int mlc_mpd_normalize(int N, const double* bins, double* counts);
/* Mpd::divide_by (mpd.cpp:71-107): PDF of mu_A/mu_B on binsR. */
int mlc_mpd_divide_by(int NA, const double* binsA, const double* countsA, int NB,
                      const double* binsB, const double* countsB, int NR, const double* binsR,
                      double* countsR);
/* trapezium (integration.cpp:3-9; integration.hpp). */
double mlc_trapezium(int N, const double* f1, const double* f2, const double* x);

#ifdef __cplusplus
}
#endif
#endif /* MLC_HOST_H */
