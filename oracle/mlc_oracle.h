/*
 * mlc_oracle.h -- FP64 CPU oracle for the microLC light-curve + chi2 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the
 * reference algorithm (gvernard/microSN, simulating_lcurves/ and chi2/),
 * used by tests/ to check results and by bench.py as the CPU baseline.  The
 * product (microsn_amd/libmlc.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Functions that
 * have no counterpart in the reference (the trajectory/magnitude chi2 and its
 * adjoint) are marked "light-curve model" and are pinned by closed forms,
 * brute force and central finite differences in tests/.
 *
 * The d/drhalf of the uniform disc is a model convention with no reference
 * counterpart (the rim's step as a one-pixel ramp); it is pinned as far as
 * the discrete disc allows, against central differences of the convolved
 * value maps over rhalf +- 1 px (within 15% in L2, correlation > 0.999).
 * Every function is pinned in tests/ (test_oracle.py, test_oracle_pins.py,
 * test_host.py, test_fft_core.py): profile values of all six shapes by numpy restatements
 * of profile.cpp and the thin disc's half-light radius, the multi-band chi2
 * bit for bit by a numpy float32 restatement of chi2_functions.cuh in both
 * index conventions, and the trajectory chi2 and its gradient by closed
 * forms on affine maps.
 */
#ifndef MLC_ORACLE_H
#define MLC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- profile shapes (profile.hpp:142-166) ---- */
enum orc_shape {
  ORC_UNIFORM = 0,       /* UniformDisc   profile.cpp:185-222 */
  ORC_GAUSSIAN = 1,      /* Gaussian      profile.cpp:228-263 */
  ORC_GAUSSIAN_HOLE = 2, /* GaussianHole  profile.cpp:269-311 */
  ORC_THERMAL_HOLE = 3,  /* ThermalHole   profile.cpp:315-354 */
  ORC_WAVY = 4,          /* Wavy          profile.cpp:358-398 */
  ORC_EXPONENTIAL = 5    /* Exponential (thin disc T~r^-3/4) profile.cpp:402-446 */
};

/* std::stof(std::to_string(x)) widened back to double: the round trip every
 * profile parameter takes in the reference (gpu_functions.cu:108,184;
 * profile.hpp:144-160). */
double orc_parity_round(double x);

/* Shape parameters derived from the half-light radius, exactly as
 * FactoryProfile::createProfileFromHalfRadius does (profile.hpp:142-166).
 * out[0] = primary parameter (R, sdev, sdev, Rin, b, sigma),
 * out[1] = secondary parameter (-, -, Rin, -, a, -).
 * Returns the (even) profile grid size n, or -1 for an unknown shape. */
int orc_profile_params(int shape, double rhalf, double pix, double extra_rin,
                       double extra_a, double* out);

/* Profile image n x n, normalised to sum 1 (profile.cpp:126-174 + shapes). */
int orc_profile_values(int shape, double rhalf, double pix, double extra_rin,
                       double extra_a, double* out, int n_cap);

/* Discretised kernel and its derivative with respect to the half-light radius
 * at a fixed grid size n (light-curve model; the grid size is the reference's
 * n(rhalf)).  dout may be NULL. */
int orc_profile_values_deriv(int shape, double rhalf, double pix,
                             double extra_rin, double extra_a, double* out,
                             double* dout, int n_cap);

/* Kernel::setKernel (magnification_map.cpp:172-194): wrap the four profile
 * quadrants into the corners of a zeroed Ny x Nx array. */
void orc_set_kernel(const double* profile, int n, int Nx, int Ny, double* kernel);

/* ---- FFT (replaces cuFFT D2Z/Z2D, gpu_functions.cu:151,157,522-549) ---- */
/* Unnormalised 1-D complex DFT, sign=-1 forward, +1 inverse.  data is
 * interleaved (re,im) of length n. */
void orc_fft1d(double* data, int n, int sign);
/* circular convolution out = IFFT(FFT(map) .* FFT(kernel)) / (Nx*Ny)
 * (gpu_functions.cu:175-200 with kernel_multiplyFFTs 571-577). */
void orc_fft_convolve(const double* map, const double* kernel, int Nx, int Ny,
                      double* out);
/* direct circular convolution, brute force, for tiny maps (test pin). */
void orc_direct_convolve(const double* map, const double* kernel, int Nx,
                         int Ny, double* out);

/* ---- locations (locations.cpp:6-25) ---- */
void orc_locations_on_grid(int Nside, int Nmap, int Noff, int* x, int* y);
void orc_locations_random(int Nloc, int Nmap, int Noff, int seed, int* x, int* y);

/* ---- sampling (kernel_sample gpu_functions.cu:604-620) ---- */
void orc_sample(const double* conv, int Nx, int k, int Nprof, int Nloc,
                const int* loc_x, const int* loc_y, double* LC, double* DLC);

/* expanding_source (gpu_functions.cu:82-225): map Ny x Nx row-major,
 * width in Einstein radii, sizes[0] must be 0. LC[Nloc*Nprof],
 * DLC[Nloc*(Nprof-1)].  Returns 0 or a negative error. */
int orc_expanding_source(const double* map, int Nx, int Ny, double width,
                         const double* sizes, int Nprof, int shape,
                         int Nloc, const int* loc_x, const int* loc_y,
                         double* LC, double* DLC);

/* ---- chi2 setup (cpu_functions.cpp:28-144) ---- */
void orc_find_indices(const double* t, const double* x, int Nd, int Nprof, int* ind);
/* Fills indA, indB, facA, facB, new_d, new_s (capacity Nd) and returns Njp;
 * *const_term receives the reference's constant term (loop quirk at
 * cpu_functions.cpp:138 reproduced when quirk_const_term != 0). */
int orc_setup_chi2(double M0, double M1, double V, double Dfac, int Nd,
                   const double* d, const double* t, const double* s,
                   int Nprof, const double* x, int quirk_const_term,
                   int* indA, int* indB, double* facA, double* facB,
                   double* new_d, double* new_s, double* const_term);

/* ---- pair mode: z, stable sort, bins (gpu_functions.cu:351-401,
 *      cpu_functions.cpp:166-189) ---- */
void orc_z(const double* LCA, const double* LCB, int Nloc, int Nprof, double* z);
/* stable argsort by '<' (std::stable_sort semantics); perm[i] = original
 * index of the i-th smallest.  -0.0 == +0.0 as in '<'; NaN keys are placed
 * after every other key, in index order (a defined order where the
 * reference's is undefined). */
void orc_stable_argsort(const double* keys, int64_t n, uint32_t* perm);
/* upper_bound counts (thrust::upper_bound + adjacent_difference,
 * gpu_functions.cu:386-396) over an ascending array. */
void orc_bin_edges(const double* sorted, int64_t n, const double* bins,
                   int Nbins, uint32_t* lower_ind, uint32_t* n_per_bin);
void orc_setup_integral(const double* LCA, const double* LCB, int Nloc,
                        int Nprof, const double* bins, int Nbins,
                        uint32_t* sorted_ind, uint32_t* lower_ind,
                        uint32_t* n_per_bin);
/* calculate_chi2_CPU (cpu_functions.cpp:213-244) */
void orc_chi2_pairs(int Njp, const int* indA, const int* indB,
                    const double* facA, const double* facB, const double* d,
                    const double* s, const double* LCA, const double* DLCA,
                    const double* LCB, const double* DLCB, int Nloc, int Nprof,
                    const uint32_t* sorted_ind, double* chi2);
/* bin_chi2 (GPU convention gpu_functions.cu:772-795 when gpu_convention!=0,
 * CPU convention cpu_functions.cpp:265-283 otherwise) */
void orc_bin_chi2(const double* chi2, const uint32_t* lower_ind,
                  const uint32_t* n_per_bin, int Nbins, int gpu_convention,
                  double* binned_chi2, double* binned_exp);

/* ---- MPDs (mpd.cpp, integration.cpp, magnification_map_gpu.cu:45-87) ---- */
void orc_mpd_bins(int Nbins, double min, double max, int log_scale, double* bins);
double orc_trapezium(int N, const double* f1, const double* f2, const double* x);
double orc_simpson(int N, const double* f1, const double* f2, const double* x);
void orc_mpd_normalize(int N, const double* bins, double* counts);
void orc_mpd_divide_by(int NA, const double* binsA, const double* countsA,
                       int NB, const double* binsB, const double* countsB,
                       int NR, const double* binsR, double* countsR);
void orc_binned_mpd(const double* map, int64_t npix, int Nbins, double min,
                    double max, double* bins, double* counts);

/* ---- GERLUMPH normalisation (magnification_map.cpp:33-38) ---- */
void orc_gerlumph_normalise(const int32_t* counts, int64_t npix, double avgmu,
                            double avgN, double k, double g, double* out);

/* ---- chi2_driver (chi2/chi2_functions.cuh:35-68,160-190), FP32 ----
 * chi2_all[N*y+x] = sum_f sum_t ((f_obs - mu1[x]/mu2[y])/df)^2/2
 * (the CPU's mu2[y]; gpu_quirk != 0 uses mu2[x] as the reference GPU kernel
 * at chi2_functions.cuh:52 does).  Summation order per pair: filter-major,
 * time-minor, one float accumulator per filter added into the pair total
 * (GPU order) when gpu_quirk, a single running float otherwise (CPU order).
 * *likelihood_f32 mimics the reference's float accumulation of exp(-chi2);
 * *likelihood_f64 is the same sum in double. */
void orc_chi2_multiband(const float* f_obs, const float* df_obs,
                        const float* mu1, const float* mu2, int N, int F, int T,
                        int gpu_quirk, float* chi2_all, float* likelihood_f32,
                        double* likelihood_f64);

/* ---- light-curve model: trajectories, magnitudes, chi2 and adjoint ----
 * See microsn_amd/csrc/mlc_traj.hip for the GPU version; this is the FP64
 * definition.
 *
 * Maps: nmaps arrays of Ny x Nx (value, d/dr) interleaved doubles.
 * Node table: for (profile p, band b): node_off[p*B+b] .. +node_cnt: global
 * node indices; node_rad[node] radius in px, node_map[node] map index.
 * Curve c: params[c*7+..] = x0, y0, theta, v, r0, w, dm; prof[c]; obs[c].
 * Obs set o, band b: epochs obs_off[o*B+b] .. obs_off[o*B+b+1] into
 * t[], mobs[], sig[].  band_scale[b]: r_b(t) = band_scale[b]*(r0 + w*t).
 * Outputs chi2[c] and grad[c*7+..] (same order as params). */
void orc_lc_chi2_grad(const double* maps, int Nx, int Ny, int B,
                      const int* node_off, const int* node_cnt,
                      const double* node_rad, const int* node_map,
                      const double* band_scale, int ncurves,
                      const double* params, const int* prof, const int* obs,
                      const int* obs_off, const double* t, const double* mobs,
                      const double* sig, double* chi2, double* grad);

/* Bilinear sample (rectGrid.cpp:349-378 convention: pixel centres at integer
 * coordinates, 0 outside the centre hull) of a plain Ny x Nx double map. */
double orc_bilinear(const double* map, int Nx, int Ny, double x, double y);

#ifdef __cplusplus
}
#endif
#endif
