/*
 * mlc.h -- C ABI of libmlc.so, the MI355X (gfx950) light-curve + chi2 path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (MLC_OK == 0) and never lets a C++ exception escape.  The text of
 * the last failure is available from mlc_last_error(ctx).
 *
 * Pointer conventions: arguments named d_* are DEVICE pointers (allocated by
 * mlc_malloc, hipMalloc or a torch tensor on the same device); every other
 * array argument is a HOST pointer.  All work is issued on the context's
 * stream; entry points that return host results synchronise that stream.
 *
 * Precision: MLC_F64 reproduces the reference arithmetic (FP64 throughout);
 * MLC_F32 is the single-precision path (FFT, taps and per-sample terms in
 * FP32, map mean / final per-curve reductions / magnitudes in FP64).
 *
 * Citations are relative to the reference tree gvernard/microSN.
 */
#ifndef MLC_H
#define MLC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLC_OK 0
#define MLC_ERR_ARG 1         /* invalid argument (checked before any launch) */
#define MLC_ERR_HIP 2         /* HIP runtime / launch failure */
#define MLC_ERR_ALLOC 3       /* device allocation failed */
#define MLC_ERR_UNSUPPORTED 4 /* size or mode outside what the kernels cover */

#define MLC_F32 0
#define MLC_F64 1

/* Profile shapes: the names accepted by FactoryProfile
 * (simulating_lcurves/include/profile.hpp:142-166). */
#define MLC_SHAPE_UNIFORM 0       /* UniformDisc  profile.cpp:185-222 */
#define MLC_SHAPE_GAUSSIAN 1      /* Gaussian     profile.cpp:228-263 */
#define MLC_SHAPE_GAUSSIAN_HOLE 2 /* GaussianHole profile.cpp:269-311 */
#define MLC_SHAPE_THERMAL_HOLE 3  /* ThermalHole  profile.cpp:315-354 */
#define MLC_SHAPE_WAVY 4          /* Wavy         profile.cpp:358-398 */
#define MLC_SHAPE_EXPONENTIAL 5   /* Exponential (thin disc) profile.cpp:402-446 */

typedef struct mlc_ctx mlc_ctx;
typedef struct mlc_map mlc_map;

/* ------------------------------------------------------------------ context
 * Replaces the reference's implicit global CUDA state (default stream,
 * __constant__ tables gpu_functions.cu:25-33, cuda_error util.cu:34-68). */
int mlc_version(void);
int mlc_ctx_create(int device, void* stream, mlc_ctx** out);
/* Frees the context's own device memory.  Maps created on it (mlc_map_create)
 * must be destroyed first: a map keeps a pointer to its context. */
void mlc_ctx_destroy(mlc_ctx* ctx);
const char* mlc_last_error(const mlc_ctx* ctx);
/* All later work of the context goes to `stream`.  Graph capture: the calls
 * of one light-curve step (mlc_map_update, mlc_map_convolve_nodes,
 * mlc_lc_chi2_grad or mlc_lc_chi2_grad_convolve, mlc_sort_f64, mlc_bin_table_device) issue no host
 * synchronisation, memset node or host copy once they have run with the same
 * sizes and tables, so a second identical step can be captured on `stream`
 * and replayed.  A captured graph refers to the context's workspace and
 * device tables: a later call that needs a larger workspace, or uploads
 * different bin edges or node jobs, invalidates it. */
int mlc_ctx_set_stream(mlc_ctx* ctx, void* stream);
/* The context's generation: a counter bumped whenever the library
 * reallocates device memory that queued or captured work may point at (the
 * workspace, a map's node tables) or rewrites one of its device tables in
 * place.  A graph captured at generation g may be replayed while
 * mlc_ctx_generation still returns g.  No reference counterpart (the
 * reference has no graph capture). */
int mlc_ctx_generation(const mlc_ctx* ctx, uint64_t* out);
/* Grows the context's workspace to at least `bytes` now (one allocation), so
 * that later calls up to that size neither reallocate it nor bump the
 * generation; used before a capture. */
int mlc_ctx_reserve(mlc_ctx* ctx, size_t bytes);
int mlc_ctx_sync(mlc_ctx* ctx);
int mlc_malloc(mlc_ctx* ctx, size_t bytes, void** d_ptr);
int mlc_free(mlc_ctx* ctx, void* d_ptr);
int mlc_memcpy_h2d(mlc_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int mlc_memcpy_d2h(mlc_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int mlc_memset(mlc_ctx* ctx, void* d_dst, int value, size_t bytes);
/* Free and total bytes of the context's device (the reference's
 * get_free_gpu_mem / get_total_gpu_mem, util.cu:125-163). */
int mlc_mem_info(mlc_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* Per-kernel HIP-event timing on the context's stream.  mlc_timing_get
 * synchronises and returns the summed duration and launch count of every
 * launch of the named kernel since the last reset. */
int mlc_timing_enable(mlc_ctx* ctx, int on);
int mlc_timing_reset(mlc_ctx* ctx);
int mlc_timing_get(mlc_ctx* ctx, const char* kernel, double* total_ms, long long* launches);
/* Newline-separated names of every kernel timed since the last reset. */
int mlc_timing_names(mlc_ctx* ctx, char* buf, size_t cap);

/* ------------------------------------------------------------------ maps
 * MagnificationMap (magnification_map.cpp:7-40): int32 ray counts ->
 * FP64 magnification, data = counts * |avgmu/avgN| / |1/((1-k)^2-g^2)|. */
int mlc_gerlumph_normalise(mlc_ctx* ctx, const int32_t* d_counts, int64_t npix,
                           double avgmu, double avgN, double k, double g,
                           double* d_map);

/* A map held for convolution (the map side of cuFFT D2Z,
 * gpu_functions.cu:151,175,522-526): the real FFT of every row of an Ny x Nx
 * row-major FP64 map, in `precision`; the full 2-D spectrum scaled by
 * 1/(Nx*Ny) (the normalisation of kernel_multiplyFFTs, gpu_functions.cu:
 * 571-577) is built from it when a call needs it.  With subtract_mean != 0
 * the FP64 mean is removed first and added back to every convolved value map
 * (the kernel sums to one).  Nx and Ny must be even, Nx/2 and Ny with prime
 * factors in {2,3,5} only. */
int mlc_map_create(mlc_ctx* ctx, int precision, const double* d_map, int Nx,
                   int Ny, int subtract_mean, mlc_map** out);
void mlc_map_destroy(mlc_map* map);
/* Re-transform a new map of the same size into an existing handle. */
int mlc_map_update(mlc_map* map, const double* d_map);

/* Source profile on its own n x n grid, normalised to sum one
 * (createProfileFromHalfRadius profile.hpp:142-166, createGrid/normalize
 * profile.cpp:143-174).  rhalf and pix are used exactly as given (the
 * reference's std::to_string/std::stof round trip is the caller's business,
 * see mlc_parity_round).  Returns the grid size in *n; d_prof/d_dprof (FP64,
 * capacity cap elements, either may be NULL) receive the profile and its
 * derivative with respect to rhalf at fixed n. */
int mlc_profile(mlc_ctx* ctx, int shape, double rhalf, double pix, double rin,
                double a, int* n, double* d_prof, double* d_dprof, int64_t cap);
/* The parameter round trip std::stof(std::to_string(x))
 * (gpu_functions.cu:108,184; profile.hpp:144-160). Host only. */
double mlc_parity_round(double x);
/* Grid size n of a profile (host only, no device work). */
int mlc_profile_size(int shape, double rhalf, double pix, double rin, double a);

/* Convolution of the map with one profile: Kernel::setKernel
 * (magnification_map.cpp:172-194) + D2Z of the kernel + kernel_multiplyFFTs +
 * Z2D (gpu_functions.cu:184-200), fused: the n x n profile is built on the
 * device, its symmetry about the pixel corner turns the column pass of the
 * 2-D convolution into a 2h-tap filter along y on the map's row spectrum,
 * and one inverse row FFT per row gives the result; nothing of size Nx*Ny is
 * staged from the host and no column FFT runs (kernels wider than the
 * staging limit fall back to fused column FFTs).
 *   d_value: Ny x Nx map in the handle's precision (float or double), or NULL
 *   d_node : Ny x Nx interleaved (value, d value / d rhalf) pairs, or NULL;
 *            stored in 4 x 4 pixel tiles: pixel (x, y) is pair number
 *            ((y/4)*(Nx/4) + x/4)*16 + (y%4)*4 + x%4 (needs Nx, Ny % 4 == 0)
 * At least one of them must be given.  A node map alone (d_value NULL, Nx a
 * power of two up to 16384) runs the filter and the row FFTs in one kernel
 * and never stores the node spectra; it agrees with the node map of a call
 * that also asks for d_value to within a few ulp of the map's scale (the
 * same sums in the same order, compiled into different kernels). */
int mlc_map_convolve(mlc_map* map, int shape, double rhalf, double pix,
                     double rin, double a, void* d_value, void* d_node);
/* Node maps of count profiles in one call: d_nodes[k] receives what
 * mlc_map_convolve(map, shape[k], rhalf[k], pix, rin[k], a[k], NULL,
 * d_nodes[k]) writes, bit for bit (rin and a may be NULL: 0 and 1).  The
 * profiles and their cosine taps of all k are built in one launch each; the
 * per-profile loop of expanding_source (gpu_functions.cu:180-210) over the
 * source sizes is the caller this serves.  Host arrays, device d_nodes[k]. */
int mlc_map_convolve_nodes(mlc_map* map, int count, const int* shape, const double* rhalf,
                           double pix, const double* rin, const double* a, void* const* d_nodes);
/* Several maps of one context and one precision in one call (many image
 * systems or lens models of a sampler, each with its own map; the reference
 * handles one map per run, gpu_functions.cu:151-175):
 *   mlc_maps_update:  mlc_map_update(maps[i], d_maps[i]) for every i;
 *   mlc_maps_convolve_nodes: mlc_map_convolve_nodes(maps[i], counts[i], ...)
 *     for every i, with map i's nodes at offsets counts[0] + ... +
 *     counts[i-1] of shape, rhalf, rin, a and d_nodes;
 * bit for bit, but the small per-map kernels (the mean and DC terms of the
 * updates; the profiles and cosine taps of the nodes) run once for all maps,
 * and when every map is at most 4096 x 4096 the maps' row transforms and
 * node maps run side by side on four more streams of the context (default
 * priority), forked from the context's stream and joined back to it by
 * events: to the caller the call is ordered on that stream as before, and
 * under graph capture the maps become parallel branches.  A map may be
 * listed more than once: it ends as the per-map calls in list order leave
 * it (updated from its last source; every listing's nodes convolved).  Host
 * arrays of nmaps (counts: of sum counts) entries. */
int mlc_maps_update(int nmaps, mlc_map* const* maps, const double* const* d_maps);
int mlc_maps_convolve_nodes(int nmaps, mlc_map* const* maps, const int* counts, const int* shape,
                            const double* rhalf, double pix, const double* rin, const double* a,
                            void* const* d_nodes);
/* The same computation structured like the reference (plain path): build
 * the full Nx x Ny kernel array on the device, forward 2-D FFT, separate
 * multiply launch, inverse 2-D FFT, separate interleave, stream-synchronised
 * between stages.  Used as the FP64 plain baseline. */
int mlc_map_convolve_plain(mlc_map* map, int shape, double rhalf, double pix,
                           double rin, double a, void* d_value, void* d_node);

/* 2-D FFTs of the library (own HIP kernels), exposed for tests:
 *   r2c: real Ny x Nx (precision) -> complex Ny x (Nx/2+1), unnormalised
 *   c2r: complex Ny x (Nx/2+1) -> real Ny x Nx, unnormalised (input kept) */
int mlc_fft_r2c(mlc_ctx* ctx, int precision, const void* d_in, void* d_out, int Nx, int Ny);
int mlc_fft_c2r(mlc_ctx* ctx, int precision, const void* d_in, void* d_out, int Nx, int Ny);

/* ------------------------------------------------------------------ sampling
 * kernel_sample (gpu_functions.cu:604-620): LC[id*Nprof+k] = conv[y*Nx+x];
 * for k > 0 also DLC[id*(Nprof-1)+k-1] = LC[id*Nprof+k] - LC[id*Nprof+k-1].
 * conv is an Ny x Nx map in `precision`; LC/DLC are FP64. */
int mlc_sample_points(mlc_ctx* ctx, int precision, const void* d_conv, int Nx,
                      int Ny, int k, int Nprof, int Nloc, const int* d_loc_x,
                      const int* d_loc_y, double* d_LC, double* d_DLC);

/* expanding_source (gpu_functions.cu:82-225; gpu_functions.hpp:11): the
 * point-source column k=0 from the raw map, then one convolution per source
 * size sizes[1..Nprof-1] (in Einstein radii; map width `width`, Rein=1),
 * with the reference's parameter round trip.  d_map is FP64 Ny x Nx. */
int mlc_expanding_source(mlc_ctx* ctx, int precision, const double* d_map,
                         int Nx, int Ny, double width, const double* sizes,
                         int Nprof, int shape, int Nloc, const int* d_loc_x,
                         const int* d_loc_y, double* d_LC, double* d_DLC);

/* ------------------------------------------------------------------ sort
 * Stable ascending argsort of FP64 keys by '<' (std::stable_sort semantics,
 * the reference's thrust::stable_sort_by_key gpu_functions.cu:370-373 and
 * its CPU twin cpu_functions.cpp:179-181).  -0.0 ties +0.0; every NaN sorts
 * after +inf, in index order.  n < 2^32.
 *   d_perm  : sorted position -> original index (required)
 *   d_rank  : original index -> sorted position (the second stable sort of
 *             gpu_functions.cu:375-377), or NULL
 *   d_sorted: keys in sorted order, or NULL */
int mlc_sort_f64(mlc_ctx* ctx, const double* d_keys, int64_t n, uint32_t* d_perm,
                 uint32_t* d_rank, double* d_sorted);
/* thrust::upper_bound + adjacent_difference (gpu_functions.cu:383-397) over
 * an ascending array: upper[i] = #{keys <= bins[i]} (NaN never counted);
 * lower_ind = [0, upper[0..Nbins-2]], n_per_bin = diff(upper).  Host out. */
int mlc_bin_edges(mlc_ctx* ctx, const double* d_sorted, int64_t n,
                  const double* bins, int Nbins, uint32_t* lower_ind,
                  uint32_t* n_per_bin);

/* ------------------------------------------------------------------ pair mode
 * setup_integral_GPU (gpu_functions.cu:351-401; gpu_functions.hpp:15):
 * z[a*Nloc+b] = LCA[a*Nprof]/LCB[b*Nprof] (kernel_z 643-658), stable sort,
 * sorted_ind = rank of every pair, bin table on `bins` (host, Nbins). */
int mlc_setup_integral(mlc_ctx* ctx, const double* d_LCA, const double* d_LCB,
                       int Nloc, int Nprof, const double* bins, int Nbins,
                       uint32_t* d_sorted_ind, uint32_t* lower_ind,
                       uint32_t* n_per_bin);
/* calculate_chi2_GPU / kernel_chi2 (gpu_functions.cu:428-456, 700-750;
 * gpu_functions.hpp:16): chi2[sorted_ind[a*Nloc+b]] =
 *   sum_i ((d_i - mA_i/mB_i)/s_i)^2,  mX_i = 1 if indX_i == -1 else
 *   LCX[.,indX_i] + facX_i*DLCX[.,indX_i].   Tables are host arrays (Njp). */
int mlc_calculate_chi2(mlc_ctx* ctx, int Njp, const int* indA, const int* indB,
                       const double* facA, const double* facB, const double* d,
                       const double* s, const uint32_t* d_sorted_ind,
                       const double* d_LCA, const double* d_DLCA,
                       const double* d_LCB, const double* d_DLCB, int Nloc,
                       int Nprof, double* d_chi2);
/* bin_chi2_GPU / kernel_bin_chi2 (gpu_functions.cu:483-503, 772-795;
 * gpu_functions.hpp:17): per-bin means of exp(-chi2/2) and chi2 over the
 * contiguous ranges [lower_ind[i], lower_ind[i]+n_per_bin[i]) of the n
 * values of d_chi2 (n = Nloc^2, the size the reference takes from its
 * Chi2SortBins); an empty bin gives (exp(-1000), 2000) as the reference GPU
 * kernel does.  Host out. */
int mlc_bin_chi2(mlc_ctx* ctx, const double* d_chi2, int64_t n,
                 const uint32_t* lower_ind, const uint32_t* n_per_bin, int Nbins,
                 double* binned_chi2, double* binned_exp);

/* The bin table and the binned means of n ascending keys without a host
 * round trip: setup_integral_GPU's upper_bound + adjacent_difference
 * (gpu_functions.cu:383-397) into d_lower / d_npb and bin_chi2_GPU
 * (gpu_functions.cu:483-503) into d_binned_chi2 / d_binned_exp, all device
 * arrays of Nbins; bins (host, Nbins right edges, non-decreasing, else
 * MLC_ERR_ARG) is copied in.  The results
 * equal mlc_bin_edges + mlc_bin_chi2 bit for bit; nothing is synchronised,
 * so a caller's loop keeps the device busy. */
int mlc_bin_table_device(mlc_ctx* ctx, const double* d_sorted, int64_t n, const double* bins,
                         int Nbins, uint32_t* d_lower, uint32_t* d_npb, double* d_binned_chi2,
                         double* d_binned_exp);

/* ------------------------------------------------------------------ MPD
 * MagnificationMap::getBinnedMpd (magnification_map_gpu.cu:45-87): counts of
 * map values v with min <= v < max on right-edge bins (host `bins`, Nbins <= 4096),
 * value v going to the first bin whose edge is >= v; counts < 1 become
 * 1/npix, the others count/npix.  Host out counts[Nbins]. */
int mlc_binned_mpd(mlc_ctx* ctx, const double* d_map, int64_t npix, int Nbins,
                   double min, double max, const double* bins, double* counts);

/* ------------------------------------------------------------------ chi2_driver
 * kernelChi2 + getChi2GPU/getChi2CPU (chi2/chi2_functions.cuh:35-190):
 *   chi2_all[N*y+x] = sum_f sum_t ((f_obs[f,t]-mu1[x,f,t]/mu2[m,f,t])/df[f,t])^2/2
 * FP32, with m = y (mode 0, the CPU twin; filters and epochs in one running
 * sum) or m = x (mode 1, the reference GPU kernel at chi2_functions.cuh:52;
 * one partial sum per filter, added in filter order).  *likelihood receives
 * sum exp(-chi2_all) over all N^2 pairs, reduced in FP64 (host out). */
int mlc_chi2_multiband(mlc_ctx* ctx, const float* d_f_obs, const float* d_df_obs,
                       const float* d_mu1, const float* d_mu2, int N, int F,
                       int T, int mode, float* d_chi2_all, double* likelihood);

/* ------------------------------------------------------------------ light curves
 * Per-curve magnitude chi2 and its adjoint along straight trajectories
 * (generalises kernel_sample gpu_functions.cu:604-620 to bilinear taps with
 * the pixel-centre convention of rectGrid.cpp:349-378, and the flux-ratio
 * chi2 of kernel_chi2 700-750 to magnitudes with offsets).
 *
 *   d_nodes : nnodes consecutive Ny x Nx maps of interleaved (mu, dmu/drad)
 *             in `precision`, 4 x 4 pixel-tiled as written by
 *             mlc_map_convolve d_node (Nx, Ny >= 4 and divisible by 4)
 *   node_off/node_cnt (host, P*B): nodes of profile p, band b are
 *             node_off[p*B+b] .. +node_cnt-1, radii ascending node_rad[]
 *   band_scale (host, B): radius of band b at time t = band_scale[b]*(r0+w*t)
 *   d_params: ncurves x 7 FP64 (x0, y0, theta, v, r0, w, dm)
 *   d_prof, d_obs: profile and observation-set index of every curve
 *   d_obs_off: (nobs*B+1) epoch offsets into the nepoch epochs d_t, d_mobs,
 *             d_sig (FP64); band b of set o is [off[o*B+b], off[o*B+b+1])
 *   out d_chi2 (ncurves, FP64) and d_grad (ncurves x 7, FP64, same order as
 *   params).  A curve leaving the pixel-centre hull or reaching mu <= 0 gets
 *   chi2 = +inf and a zero gradient.
 * Between radius nodes mu is the cubic Hermite interpolant of (mu, dmu/drad);
 * below the first / above the last node it is the first-order Taylor step.
 * Under graph capture the band launches of different profiles become
 * parallel branches on four more streams of the context (default priority,
 * created by an earlier call that was not captured); issued eagerly they run
 * in order on the context's stream.  Results are the same bits either way. */
int mlc_lc_chi2_grad(mlc_ctx* ctx, int precision, const void* d_nodes,
                     int nnodes, int Nx, int Ny, int P, int B,
                     const int* node_off, const int* node_cnt,
                     const double* node_rad, const double* band_scale,
                     int ncurves, const double* d_params, const int* d_prof,
                     const int* d_obs, int nobs, const int* d_obs_off,
                     int nepoch, const double* d_t, const double* d_mobs,
                     const double* d_sig, double* d_chi2, double* d_grad);
/* The node maps and the chi2 of one step in one call:
 *   mlc_map_convolve_nodes(map, nnodes, node_shape, node_rad, pix, NULL, NULL,
 *                          {d_nodes + k * Nx*Ny pairs})
 * followed by mlc_lc_chi2_grad(ctx of map, precision of map, d_nodes, ...),
 * with bit-identical node maps, chi2 and gradient, in one call (it replaces
 * the per-profile convolve-then-sample loop of expanding_source,
 * gpu_functions.cu:180-225; every node map is written before the first
 * band launch).  d_nodes: nnodes consecutive node maps (the
 * layout mlc_map_convolve writes to d_node); node_shape: host array of
 * nnodes MLC_SHAPE_* values; node_rad doubles as each node's half-light
 * radius (in units of pix) for the convolution.  Every node map is written,
 * including those no band reads and when ncurves is 0.  The curve
 * partition runs on a second stream of the context (default priority,
 * created on the first call that is not being captured) beside the node
 * kernels, forked from and joined back into the context's stream by events
 * inside the call.  Under graph capture the node kernels too (when every
 * node map comes from the fused kernel alone) are parallel branches on the
 * four streams of mlc_lc_chi2_grad's band launches. */
int mlc_lc_chi2_grad_convolve(mlc_map* map, const int* node_shape, double pix,
                              void* d_nodes, int nnodes, int P, int B,
                              const int* node_off, const int* node_cnt,
                              const double* node_rad, const double* band_scale,
                              int ncurves, const double* d_params,
                              const int* d_prof, const int* d_obs, int nobs,
                              const int* d_obs_off, int nepoch, const double* d_t,
                              const double* d_mobs, const double* d_sig,
                              double* d_chi2, double* d_grad);
/* Several maps in one call (batched light curves): curve c lies on map
 * d_map[c] (0 <= d_map[c] < nmaps; another value gives chi2 = +inf), has
 * profile d_prof[c] and observation set d_obs[c].  Map m is Ny[m] x Nx[m]
 * (host arrays; multiples of 4).  d_nodes (host array of nnodes device
 * pointers) gives node map k, tiled like mlc_map_convolve d_node, in
 * `precision`; the nodes of (map m, profile p, band b) are node_off[j] ..
 * node_off[j] + node_cnt[j] - 1, j = (m*P + p)*B + b, radii node_rad
 * ascending (1..16 per job), and a node map serves the jobs of one map
 * only.  Every other argument means what it means for mlc_lc_chi2_grad.
 * All bands of all maps run in one persistent launch that takes groups of
 * curves from a per-XCD work queue (each group on one (map, profile), its
 * bands in order), after a counting partition of the curves by (map,
 * profile, map cell).  Each curve's chi2 and gradient are those
 * mlc_lc_chi2_grad gives on its map alone with the same launch shape (bit
 * for bit), so maps of different sizes and curves with uneven epoch counts
 * share one launch.  Generalises the per-profile serial loop of
 * expanding_source (gpu_functions.cu:180-210) to many maps. */
int mlc_lc_chi2_grad_batch(mlc_ctx* ctx, int precision, int nmaps, const int* Nx,
                           const int* Ny, int nnodes, void* const* d_nodes,
                           const double* node_rad, int P, int B, const int* node_off,
                           const int* node_cnt, const double* band_scale, int ncurves,
                           const double* d_params, const int* d_map, const int* d_prof,
                           const int* d_obs, int nobs, const int* d_obs_off, int nepoch,
                           const double* d_t, const double* d_mobs, const double* d_sig,
                           double* d_chi2, double* d_grad);

/* Launch-shape tuner of the band kernels (mlc_lc_chi2_grad and its
 * variants, mlc_lc_chi2_grad_batch).  A shape is the lanes per curve (8, 16
 * or 32) and the grid as a multiple of the workgroups the CUs hold at once.
 * MLC_TUNE_AUTO (the default): the first call with a new key (kernel kind,
 * precision, log2 of the mean epochs per curve and band, log2 of the
 * largest map's pixels) times the six candidates on its own inputs and keeps
 * the fastest for the context; a call under graph capture never tunes (it
 * uses a kept shape, else 16 lanes and one grid fill).  MLC_TUNE_OFF: always
 * 16 lanes, one grid fill.  MLC_TUNE_FORCE: the shape of mlc_tune_force.
 * Shapes differ only in the order of the FP64 per-band sums.  No reference
 * counterpart (its launch geometry is fixed, gpu_functions.cu:88-91). */
#define MLC_TUNE_OFF 0
#define MLC_TUNE_AUTO 1
#define MLC_TUNE_FORCE 2
#define MLC_TUNE_KIND_BANDS 0 /* per-(profile, band) launches */
#define MLC_TUNE_KIND_BATCH 1 /* the persistent batched launch */
/* Force flag: the multi-node band kernel also for single-node bands. */
#define MLC_TUNE_GENERAL_NODES 1
/* Shape flag of the batched kernel: one launch per band (all maps) instead
 * of every band in one persistent launch; a candidate of the tuner when a
 * batched call has more than one band (it keeps one band's node maps in the
 * caches at a time). */
#define MLC_TUNE_BATCH_PER_BAND 2
int mlc_tune_mode(mlc_ctx* ctx, int mode);
int mlc_tune_force(mlc_ctx* ctx, int lanes, int grid_multiple, int flags);
/* The kept shape for a key (lanes = grid_multiple = 0 when none is kept;
 * *flags receives MLC_TUNE_BATCH_PER_BAND when that form was kept). */
int mlc_tune_lookup(mlc_ctx* ctx, int kind, int precision, double epochs_per_band,
                    double pixels, int* lanes, int* grid_multiple, int* flags);
/* Kept shapes to / from a text file (loading adds to the kept shapes). */
int mlc_tune_save(mlc_ctx* ctx, const char* path);
int mlc_tune_load(mlc_ctx* ctx, const char* path);

/* mlc_lc_chi2_grad without any device allocation: the same kernels and the
 * same results bit for bit, but the call's scratch (curve records, partition
 * counts, per-curve accumulators, epoch records: about 136 bytes per curve
 * plus 32 per epoch) comes from a static device array of 64 MiB per GPU
 * (held by each GPU the library runs on until the process ends), not from
 * the context's growable workspace.  A call whose scratch would exceed
 * 64 MiB fails with MLC_ERR_ARG before any launch; nothing calls hipMalloc
 * or hipFree.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two contexts, two host threads) race and
 * must be serialised by the caller.  No counterpart in the reference; it is
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). */
int mlc_lc_chi2_grad_noalloc(mlc_ctx* ctx, int precision, const void* d_nodes,
                             int nnodes, int Nx, int Ny, int P, int B,
                             const int* node_off, const int* node_cnt,
                             const double* node_rad, const double* band_scale,
                             int ncurves, const double* d_params, const int* d_prof,
                             const int* d_obs, int nobs, const int* d_obs_off,
                             int nepoch, const double* d_t, const double* d_mobs,
                             const double* d_sig, double* d_chi2, double* d_grad);
/* Same result, structured like the reference's plain pipeline: one launch
 * writes per-sample (mu, dmu/dx, dmu/dy, dmu/drad) to device memory, the
 * stream is synchronised, a second launch reduces chi2 and gradient.  The
 * FP64 plain baseline. */
int mlc_lc_chi2_grad_plain(mlc_ctx* ctx, int precision, const void* d_nodes,
                           int nnodes, int Nx, int Ny, int P, int B,
                           const int* node_off, const int* node_cnt,
                           const double* node_rad, const double* band_scale,
                           int ncurves, const double* d_params,
                           const int* d_prof, const int* d_obs, int nobs,
                           const int* d_obs_off, int nepoch, const double* d_t,
                           const double* d_mobs, const double* d_sig,
                           double* d_chi2, double* d_grad);

#ifdef __cplusplus
}
#endif
#endif /* MLC_H */
