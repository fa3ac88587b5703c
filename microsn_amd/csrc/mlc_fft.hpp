// mlc_fft.hpp -- the map handle and the precision-templated FFT entry points
// shared by the C ABI (mlc_fft.hip) and the kernel translation units
// (mlc_fft_f32.hip, mlc_fft_f64.hip).
#pragma once

#include <functional>
#include <vector>

#include "mlc_fft_core.hpp"
#include "mlc_internal.hpp"
#include "mlc_profile.hpp"

struct mlc_map {
  mlc_ctx* ctx = nullptr;
  int precision = MLC_F32;
  int Nx = 0, Ny = 0, W = 0;
  int ldr = 0;             // row stride of rows: W rounded up to even (16-byte aligned column pairs)
  int subtract_mean = 0;
  void* rows = nullptr;    // Ny x ldr complex (precision): R2C of every row of map - mean
  void* fmap = nullptr;    // Ny x W complex: 2-D spectrum / (Nx*Ny), built on first use
  bool full_ready = false; // fmap matches rows
  double* mean = nullptr;  // device FP64: map mean (0 when not subtracted)
  void* g0 = nullptr;      // Ny x W complex scratch (value column output)
  void* g1 = nullptr;      // Ny x W complex scratch (derivative column output)
  void* ntab = nullptr;    // node profiles and taps of the last mlc_map_convolve_nodes call
  size_t ntab_bytes = 0;
  mlc_ctx::ConstTable jobs; // device copy of the node jobs of that call
};

namespace mlc {

// Throws unless Nx is even, Nx/2 and Ny factor over {2,3,5} and the offsets
// fit the kernels' 32-bit indexing.
void fft_check_dims(int Nx, int Ny);

// n x n profile and d/drhalf, normalised, into device FP64 arrays.
void fft_profile(mlc_ctx* ctx, const ProfileParams& P, double* prof, double* dprof);

// Row spectrum of d_map - mean into m->rows (and the FP64 mean into m->mean).
template <class T>
void fft_map_transform(mlc_map* m, const double* d_map);
// Fused convolution (see mlc_map_convolve).
template <class T>
void fft_convolve_fused(mlc_map* m, const ProfileParams& P, void* d_value, void* d_node);
// Node maps of several profiles (see mlc_map_convolve_nodes).  Without a
// driver the nodes are convolved in order; with one, drive(conv, independent)
// is called once after the profiles and taps of all nodes are built, and
// conv(k) convolves node k when the driver calls it (each node once).
// independent: no two nodes share a scratch buffer, so conv(k) of different
// nodes may run on different streams.
using NodeDriver = std::function<void(const std::function<void(int)>& conv, bool independent)>;
template <class T>
void fft_convolve_nodes(mlc_map* m, const std::vector<ProfileParams>& Ps, void* const* d_nodes,
                        const NodeDriver* drive = nullptr);
// Several maps of one context (see mlc_maps_update, mlc_maps_convolve_nodes):
// the same kernels and sums as the per-map calls, with the small per-map
// kernels of all maps in one launch each.  Ps[i] are the nodes of map i;
// d_nodes lists the node maps of map 0, then map 1, ...
template <class T>
void fft_maps_transform(const std::vector<mlc_map*>& maps, const double* const* d_maps);
template <class T>
void fft_maps_convolve_nodes(const std::vector<mlc_map*>& maps, const std::vector<std::vector<ProfileParams>>& Ps,
                             void* const* d_nodes);
// Reference-structured convolution (see mlc_map_convolve_plain).
template <class T>
void fft_convolve_plain(mlc_map* m, const ProfileParams& P, void* d_value, void* d_node);
// Unnormalised 2-D real transforms; fft_c2r_2d overwrites `work`.
template <class T>
void fft_r2c_2d(mlc_ctx* ctx, const T* in, cpx<T>* out, int Nx, int Ny);
template <class T>
void fft_c2r_2d(mlc_ctx* ctx, cpx<T>* work, T* out, int Nx, int Ny);

extern template void fft_map_transform<float>(mlc_map*, const double*);
extern template void fft_map_transform<double>(mlc_map*, const double*);
extern template void fft_convolve_fused<float>(mlc_map*, const ProfileParams&, void*, void*);
extern template void fft_convolve_fused<double>(mlc_map*, const ProfileParams&, void*, void*);
extern template void fft_convolve_nodes<float>(mlc_map*, const std::vector<ProfileParams>&, void* const*,
                                               const NodeDriver*);
extern template void fft_convolve_nodes<double>(mlc_map*, const std::vector<ProfileParams>&, void* const*,
                                                const NodeDriver*);
extern template void fft_maps_transform<float>(const std::vector<mlc_map*>&, const double* const*);
extern template void fft_maps_transform<double>(const std::vector<mlc_map*>&, const double* const*);
extern template void fft_maps_convolve_nodes<float>(const std::vector<mlc_map*>&,
                                                    const std::vector<std::vector<ProfileParams>>&, void* const*);
extern template void fft_maps_convolve_nodes<double>(const std::vector<mlc_map*>&,
                                                     const std::vector<std::vector<ProfileParams>>&, void* const*);
extern template void fft_convolve_plain<float>(mlc_map*, const ProfileParams&, void*, void*);
extern template void fft_convolve_plain<double>(mlc_map*, const ProfileParams&, void*, void*);
extern template void fft_r2c_2d<float>(mlc_ctx*, const float*, cpx<float>*, int, int);
extern template void fft_r2c_2d<double>(mlc_ctx*, const double*, cpx<double>*, int, int);
extern template void fft_c2r_2d<float>(mlc_ctx*, cpx<float>*, float*, int, int);
extern template void fft_c2r_2d<double>(mlc_ctx*, cpx<double>*, double*, int, int);

}  // namespace mlc
