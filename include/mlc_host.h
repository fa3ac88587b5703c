/*
 * mlc_host.h -- host-only entry points of libmlc.so: the driver-side steps of
 * simulating_lcurves that prepare the device path's inputs.  Same C ABI
 * conventions as mlc.h (int status, 0 = success; host arrays only).
 * Citations are relative to the reference tree gvernard/microSN.
 */
#ifndef MLC_HOST_H
#define MLC_HOST_H

#include <stdint.h>

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

/* GERLUMPH map directory (magnification_map.cpp:7-40): <dir>/mapmeta.dat
 * ("avgmu avgN / Nx / width / k g s") and <dir>/map.bin (Nx*Nx native int32
 * ray counts).  Fills the metadata; with counts != NULL (capacity cap >=
 * Nx*Nx) also reads the counts.  Returns 1 for bad arguments, 2 for missing
 * or malformed metadata, 3 for a missing or short map.bin (the reference does
 * not check either file). */
int mlc_gerlumph_read(const char* dir, double* avgmu, double* avgN, int* Nx, double* width,
                      double* k, double* g, double* s, int32_t* counts, int64_t cap);
/* Writes a map directory in the same format (no reference counterpart; for
 * synthetic maps). */
int mlc_gerlumph_write(const char* dir, double avgmu, double avgN, int Nx, double width, double k,
                       double g, double s, const int32_t* counts);
/* Image::writeImageFITS (image.cpp:20-62): every sampling-th pixel of an
 * Ny x Nx FP64 image as FLOAT_IMG, rows flipped, in the primary HDU and in an
 * image extension "NEW-EXTENSION". */
int mlc_write_fits(const char* path, const double* data, int Nx, int Ny, int sampling);
/* FITS image input, the optional map input (the reference reads maps only
 * as GERLUMPH map.bin, magnification_map.cpp:7-40): HDU hdu (0 = primary)
 * of a basic FITS file, 2-D, BITPIX 8/16/32/64/-32/-64, BSCALE/BZERO
 * applied, rows flipped back to the orientation mlc_write_fits was given.
 * Sets *Nx, *Ny; data (capacity cap) receives Ny x Nx FP64.  Returns 1 for
 * bad arguments, 2 for a missing file, a malformed header or no such 2-D
 * HDU, 3 when cap < Nx*Ny (the sizes are set). */
int mlc_read_fits(const char* path, int hdu, double* data, int64_t cap, int* Nx, int* Ny);
/* Image::writeImageBIN (image.cpp:4-18): the same samples as native FP64. */
int mlc_write_bin(const char* path, const double* data, int Nx, int Ny, int sampling);

/* BaseProfile size laws (profile.cpp:86-124) and their throughput-weighted
 * forms over a filter curve (profile.cpp:55-84, trapezoid rule). */
double mlc_size_parametric(double r0, double l0, double nu, double lrest);
double mlc_size_ss(double mbh, double fedd, double eta, double lrest);
double mlc_size_parametric_band(double r0, double l0, double nu, const double* wavelength,
                                const double* throughput, int n);
double mlc_size_ss_band(double mbh, double fedd, double eta, const double* wavelength,
                        const double* throughput, int n);
double mlc_size_vector_band(const double* wavelength, const double* throughput, int n);

#ifdef __cplusplus
}
#endif
#endif /* MLC_HOST_H */
