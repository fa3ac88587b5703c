// mlc_fft_impl.hpp -- the library's own 2-D FFT and the fused map x profile
// convolution (replaces cuFFT D2Z/Z2D + kernel_multiplyFFTs +
// Kernel::setKernel, gpu_functions.cu:151-200, 522-577;
// magnification_map.cpp:172-194).
//
// Layout.  A real Ny x Nx map is row-major; its half spectrum is Ny rows of
// W = Nx/2+1 complex values (the cuFFT D2Z layout).  Every transform runs
// inside LDS:
//   row pass    one workgroup per row: the real row of length Nx is folded
//               into a complex sequence of length M = Nx/2 (even/odd
//               samples), transformed, and unfolded (mlc_fft_core.hpp);
//   column pass one workgroup per CB adjacent columns of the half spectrum,
//               each column transformed by 1024/CB threads; global reads
//               and writes move CB*8 (or CB*16) contiguous bytes per row and
//               workgroups are numbered so that neighbouring column groups
//               run on the same XCD (shared L2).
// Per source profile (mlc_map_convolve) nothing of size Nx*Ny is staged from
// the host: the n x n profile is evaluated on the device, its wrapped kernel
// is transformed along x directly (n nonzero rows, n nonzero columns), and a
// single column kernel does forward FFT of the kernel column -> multiply by
// the map spectrum column -> inverse FFT, for the value and the d/drhalf
// kernel.  A row pass then writes the real maps.
#pragma once

#include <algorithm>
#include <cstring>
#include <cmath>
#include <type_traits>
#include <vector>

#include "mlc_fft_core.hpp"
#include "mlc_internal.hpp"
#include "mlc_ctfft.hpp"
#include "mlc_fft.hpp"
#include "mlc_profile.hpp"
#include "mlc_tiles.hpp"

using namespace mlc;

namespace {

constexpr int kColThreads = 1024;
constexpr size_t kLdsBytes = 160 * 1024;

// ------------------------------------------------------------------ device FFT
// Twiddles live in LDS as W_L^m, m < QL: a quarter wave (QL = L/4) when
// 4 | L, else the full table (QL = L); QuarterTw reconstructs the rest.
__host__ __device__ inline int tw_len(int L) { return (L % 4 == 0) ? L / 4 : L; }

template <class T, int R, int G, int EPT, bool INV, class TW, class PAD>
__device__ __forceinline__ void dev_stage(cpx<T>* buf, int L, int Ns, const TW& tw, PAD pad, int g) {
  constexpr int BMAX = (EPT + R - 1) / R;
  cpx<T> v[BMAX][R];
  const int nb = L / R;
#pragma unroll
  for (int i = 0; i < BMAX; ++i) {
    int j = g + i * G;
    if (j < nb) stage_load<T, R, INV>(buf, L, Ns, j, tw, v[i], pad);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < BMAX; ++i) {
    int j = g + i * G;
    if (j < nb) stage_store<T, R>(buf, Ns, j, v[i], pad);
  }
  __syncthreads();
}

// All stages of `plan` on buf by the G threads of one group (g = index in
// the group).  Every thread of the workgroup must call this: each stage
// synchronises the whole workgroup, and all groups run the same plan.
template <class T, int G, int EPT, bool INV, class TW, class PAD>
__device__ void dev_fft(cpx<T>* buf, const Plan& plan, const TW& tw, PAD pad, int g) {
  int Ns = 1;
  for (int s = 0; s < plan.n; ++s) {
    const int R = plan.r[s];
    // radix 16 only where the register budget allows it (plans for
    // 1024-thread column kernels never contain it)
    if (G <= 256 && R == 16) dev_stage<T, 16, G, EPT, INV>(buf, plan.L, Ns, tw, pad, g);
    else if (R == 8) dev_stage<T, 8, G, EPT, INV>(buf, plan.L, Ns, tw, pad, g);
    else if (R == 4) dev_stage<T, 4, G, EPT, INV>(buf, plan.L, Ns, tw, pad, g);
    else if (R == 2) dev_stage<T, 2, G, EPT, INV>(buf, plan.L, Ns, tw, pad, g);
    else if (R == 5) dev_stage<T, 5, G, EPT, INV>(buf, plan.L, Ns, tw, pad, g);
    else dev_stage<T, 3, G, EPT, INV>(buf, plan.L, Ns, tw, pad, g);
    Ns *= R;
  }
}

// The same transform with the plan fixed at compile time (length L known to
// the compiler): the stages are those make_plan(L, 8) gives, in the same
// order with the same arithmetic, so results are bit-identical to dev_fft,
// but every index division and modulo by the stage constants folds into
// multiplies and shifts.
struct FixedPlan {
  int n;
  int r[24];
};

constexpr FixedPlan fixed_plan(int L) {
  FixedPlan p{0, {}};
  int m = L;
  while (m % 8 == 0) { p.r[p.n++] = 8; m /= 8; }
  while (m % 4 == 0) { p.r[p.n++] = 4; m /= 4; }
  while (m % 2 == 0) { p.r[p.n++] = 2; m /= 2; }
  while (m % 5 == 0) { p.r[p.n++] = 5; m /= 5; }
  while (m % 3 == 0) { p.r[p.n++] = 3; m /= 3; }
  return p;
}

template <class T, int L, int G, int EPT, bool INV, int S, int NS, class TW, class PAD>
__device__ __forceinline__ void fixed_fft(cpx<T>* buf, const TW& tw, PAD pad, int g) {
  constexpr FixedPlan P = fixed_plan(L);
  if constexpr (S < P.n) {
    constexpr int R = P.r[S];
    dev_stage<T, R, G, EPT, INV>(buf, L, NS, tw, pad, g);
    fixed_fft<T, L, G, EPT, INV, S + 1, NS * R>(buf, tw, pad, g);
  }
}

// Transform of a row of runtime length (ML == 0, plan `plan`) or of the
// compile-time length ML.
template <class T, int ML, int G, bool INV, class TW, class PAD>
__device__ __forceinline__ void row_fft(cpx<T>* buf, const Plan& plan, const TW& tw, PAD pad, int g) {
  if constexpr (ML > 0) fixed_fft<T, ML, G, 16, INV, 0, 1>(buf, tw, pad, g);
  else dev_fft<T, G, 16, INV>(buf, plan, tw, pad, g);
}

// A store into a node map by the fused kernels, which write 64 or 128
// contiguous bytes of a tile per workgroup.  Non-temporal: node maps are
// written once per step and read by a later kernel, so they should not push
// the row spectrum, which the y-filter reads many times over, out of the
// caches (the 4096^2 bench step took 2-3% less time with these).  The row
// kernels keep plain stores: they write 32-byte pieces of a tile from
// different workgroups, which the L2 has to merge (non-temporal, the
// 10000-wide node rows took 1.7 times as long).
template <class V>
__device__ __forceinline__ void node_store(V* p, const V& v) {
  __builtin_nontemporal_store(v, p);
}

__device__ __forceinline__ int wrap_row(int y, int h, int N) {
  int s = y + h;
  return s >= N ? s - N : s;
}

// ------------------------------------------------------------------ row kernels
// LDS of a row kernel: [twiddles W_M^m, m < tw_len(M)] [row buffer(s)].
// Forward real FFT of one row per workgroup.  in: Ny x Nx of TIN; out: Ny
// rows of W = Nx/2 + 1 bins, row stride ldo.
// With shift != NULL the FP64 value *shift is subtracted before rounding to T.
// With rowsum != NULL the row's own FP64 mean is subtracted instead and its
// FP64 sum goes to rowsum[y] (k_row_dc_fix then sets the DC terms).
// ML > 0: the half length M = Nx/2 is the compile-time constant ML.
template <class T, class TIN, int G, int ML>
__global__ __launch_bounds__(G) void k_row_r2c(const TIN* __restrict__ in, cpx<T>* __restrict__ out,
                                               int ldo, int Nx, const double* __restrict__ shift,
                                               double* __restrict__ rowsum, Plan plan,
                                               const cpx<T>* __restrict__ twM,
                                               const cpx<T>* __restrict__ twN) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int y = blockIdx.x, g = threadIdx.x, M = ML > 0 ? ML : Nx >> 1, QL = tw_len(M);
  if constexpr (ML > 0) Nx = 2 * ML;
  cpx<T>* q = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* buf = q + QL;
  const NoPad P;
  for (int i = g; i < QL; i += G) q[i] = twM[i];
  const cpx<TIN>* row = reinterpret_cast<const cpx<TIN>*>(in + (size_t)y * Nx);
  double sh = shift ? *shift : 0.0;
  // the row in registers (M <= 16 G for every row group)
  constexpr int kPer = 16;
  cpx<TIN> v[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int n = g + i * G;
    if (n < M) v[i] = row[n];
  }
  if (rowsum) {
    // the row buffer, not yet filled, holds the per-wave partial sums (no
    // static LDS: the widest rows use all of it)
    double* red = reinterpret_cast<double*>(buf);
    double a = 0.0;
#pragma unroll
    for (int i = 0; i < kPer; ++i)
      if (g + i * G < M) a += (double)v[i].x + (double)v[i].y;
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((g & 63) == 0) red[g >> 6] = a;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < G / 64; ++w) tot += red[w];
    __syncthreads();  // every wave has read the sums before buf is filled
    if (g == 0) rowsum[y] = tot;
    sh = tot / (double)Nx;
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int n = g + i * G;
    if (n < M) buf[P(n)] = mk<T>((T)((double)v[i].x - sh), (T)((double)v[i].y - sh));
  }
  __syncthreads();
  row_fft<T, ML, G, false>(buf, plan, QuarterTw<T>{q, QL}, P, g);
  cpx<T>* o = out + (size_t)y * ldo;
  for (int k = g; k <= (M >> 1); k += G) {
    cpx<T> Xk, Xm;
    r2c_post(buf[P(k == 0 ? 0 : k)], buf[P(k == 0 ? 0 : M - k)], twN[k], &Xk, &Xm);
    o[k] = Xk;
    if (M - k != k) o[M - k] = Xm;
  }
}

// After k_row_r2c with row sums: the map mean (sum of the row sums in a
// fixed order / (Nx Ny)) into *mean and the DC term of every row spectrum,
// rowsum[y] - Nx mean, the FFT of the row minus the map mean would give.
template <class T>
__device__ __forceinline__ void row_dc_fix_block(const double* __restrict__ rowsum, int Nx, int Ny,
                                                 cpx<T>* __restrict__ out, int ldo, double* __restrict__ mean) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  double a = 0.0;
  for (int y = t; y < Ny; y += 1024) a += rowsum[y];
  red[t] = a;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double mu = red[0] / ((double)Nx * (double)Ny);
  if (t == 0) *mean = mu;
  for (int y = t; y < Ny; y += 1024) out[(size_t)y * ldo] = mk<T>((T)(rowsum[y] - (double)Nx * mu), (T)0);
}

template <class T>
__global__ __launch_bounds__(1024) void k_row_dc_fix(const double* __restrict__ rowsum, int Nx, int Ny,
                                                     cpx<T>* __restrict__ out, int ldo,
                                                     double* __restrict__ mean) {
  row_dc_fix_block<T>(rowsum, Nx, Ny, out, ldo, mean);
}

// The DC step of several maps, one workgroup per map (the batched update).
struct DcJob {
  const double* rowsum;
  void* out;
  double* mean;
  int Nx, Ny, ldo;
};
template <class T>
__global__ __launch_bounds__(1024) void k_row_dc_fix_many(const DcJob* __restrict__ jobs) {
  const DcJob& J = jobs[blockIdx.x];
  row_dc_fix_block<T>(J.rowsum, J.Nx, J.Ny, static_cast<cpx<T>*>(J.out), J.ldo, J.mean);
}

// Loads the folded inverse input of one row into buf (see c2r_pre).
template <class T, int G>
__device__ __forceinline__ void c2r_fill(cpx<T>* buf, const cpx<T>* __restrict__ X, int M,
                                         const cpx<T>* __restrict__ twN, int g) {
  const NoPad P;
  for (int k = g; k <= (M >> 1); k += G) {
    cpx<T> Zk, Zm;
    c2r_pre(X[k], X[M - k], twN[k], &Zk, &Zm);
    buf[P(k)] = Zk;
    if (k > 0 && M - k != k) buf[P(M - k)] = Zm;
  }
}

// Inverse real FFT of one row per workgroup (unnormalised).
// MODE 0: out is T[Ny][Nx] = x;  MODE 1: out is T[Ny][Nx] = x + *shift;
// MODE 2: out is the node map (x0 + *shift, x1) of the spectra in0, in1 in
// 4 x 4 pixel tiles (mlc_tiles.hpp); rows are numbered XCD-contiguously so
// the four rows of a tile are written through one L2.  MODE 3 / MODE 4:
// only the first (x0 + *shift) / second (x1) member of every node pair, from
// in0 alone, for rows too long to hold two buffers in LDS.
// ML > 0: the half length M = Nx/2 is the compile-time constant ML.
// GTW: the FFT's twiddles are read from the global table (through the
// caches) instead of an LDS copy, so that the row buffers alone fill the LDS
// and two workgroups fit on a CU (the 10000-wide node rows: 2 x 40 KB).
template <class T, int G, int MODE, int ML, bool GTW = false>
__global__ __launch_bounds__(G) void k_row_c2r(const cpx<T>* __restrict__ in0,
                                               const cpx<T>* __restrict__ in1, T* __restrict__ out,
                                               int Nx, int Ny, const double* __restrict__ shift,
                                               Plan plan, const cpx<T>* __restrict__ twM,
                                               const cpx<T>* __restrict__ twN) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int y = MODE >= 2 ? xcd_block(Ny) : (int)blockIdx.x;
  if (y >= Ny) return;
  if constexpr (ML > 0) Nx = 2 * ML;
  const int g = threadIdx.x, M = Nx >> 1, W = M + 1, QL = tw_len(M);
  const cpx<T>* q;
  cpx<T>* buf0;
  if constexpr (GTW) {
    q = twM;
    buf0 = reinterpret_cast<cpx<T>*>(smem);
  } else {
    cpx<T>* ql = reinterpret_cast<cpx<T>*>(smem);
    for (int i = g; i < QL; i += G) ql[i] = twM[i];
    q = ql;
    buf0 = ql + QL;
  }
  const NoPad P;
  c2r_fill<T, G>(buf0, in0 + (size_t)y * W, M, twN, g);
  cpx<T>* buf1 = buf0 + M;
  if constexpr (MODE == 2) c2r_fill<T, G>(buf1, in1 + (size_t)y * W, M, twN, g);
  __syncthreads();
  const QuarterTw<T> tw{q, QL};
  row_fft<T, ML, G, true>(buf0, plan, tw, P, g);
  const T sh = (MODE != 0 && shift) ? (T)*shift : (T)0;
  if constexpr (MODE == 2) {
    row_fft<T, ML, G, true>(buf1, plan, tw, P, g);
    // (value, deriv) pairs of pixels 2n and 2n+1: adjacent in their tile
    // row, one 2-pair store
    typedef T v4 __attribute__((ext_vector_type(4)));
    cpx<T>* o = reinterpret_cast<cpx<T>*>(out);
    for (int n = g; n < M; n += G) {
      cpx<T> a = buf0[P(n)], b = buf1[P(n)];
      const v4 val = {a.x + sh, b.x, a.y + sh, b.y};
      *reinterpret_cast<v4*>(o + tile_offset(2 * n, y, Nx)) = val;
    }
  } else if constexpr (MODE == 3 || MODE == 4) {
    cpx<T>* o = reinterpret_cast<cpx<T>*>(out);
    for (int n = g; n < M; n += G) {
      cpx<T> a = buf0[P(n)];
      cpx<T>* dst = o + tile_offset(2 * n, y, Nx);
      if constexpr (MODE == 3) {
        dst[0].x = a.x + sh;
        dst[1].x = a.y + sh;
      } else {
        dst[0].y = a.x;
        dst[1].y = a.y;
      }
    }
  } else {
    cpx<T>* o = reinterpret_cast<cpx<T>*>(out + (size_t)y * Nx);
    for (int n = g; n < M; n += G) {
      cpx<T> a = buf0[P(n)];
      o[n] = mk<T>(a.x + sh, a.y + sh);
    }
  }
}

// The 10000-wide node rows (MODE 2 of k_row_c2r, ML = 5000) as a
// persistent kernel, one workgroup per CU: each workgroup transforms rows y,
// y + s, ... of its XCD's contiguous range (s = workgroups on that XCD) and
// loads the next row's spectra into registers while it transforms the
// current one (the FFT reads its twiddles from an LDS copy, so no wait on
// global loads drains them early).  One row per workgroup, two workgroups per
// CU with the twiddles read from the global table, loaded, transformed and
// stored in lockstep, each phase leaving the other units idle.  Same
// arithmetic as k_row_c2r MODE 2 (c2r_pre, row_fft, the 2-pair stores), so
// the same bits.
template <class T, int G, int ML>
__global__ __launch_bounds__(G) void k_row_c2r_node_pf(const cpx<T>* __restrict__ in0,
                                                       const cpx<T>* __restrict__ in1, T* __restrict__ out,
                                                       int Ny, const double* __restrict__ shift, Plan plan,
                                                       const cpx<T>* __restrict__ twM,
                                                       const cpx<T>* __restrict__ twN) {
  constexpr int M = ML, Nx = 2 * ML, W = M + 1, KPT = (M / 2 + G) / G, QL = ML % 4 == 0 ? ML / 4 : ML;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* ql = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* buf0 = ql + QL;
  cpx<T>* buf1 = buf0 + M;
  const int g = threadIdx.x, x = blockIdx.x & 7, nwx = gridDim.x >> 3;
  const int per = (Ny + 7) >> 3, ya = x * per, yb = min(Ny, ya + per);
  int y = ya + (int)(blockIdx.x >> 3);
  if (y >= yb) return;  // uniform per workgroup
  cpx<T> tw[KPT], a0[KPT], b0[KPT], a1[KPT], b1[KPT];  // twN[k]; X0[k], X0[M-k], X1[k], X1[M-k]
  auto load = [&](int yy) {
    const cpx<T>* r0 = in0 + (size_t)yy * W;
    const cpx<T>* r1 = in1 + (size_t)yy * W;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const int k = g + j * G;
      if (k <= M / 2) {
        a0[j] = r0[k];
        b0[j] = r0[M - k];
        a1[j] = r1[k];
        b1[j] = r1[M - k];
      }
    }
  };
  load(y);
  for (int i = g; i < QL; i += G) ql[i] = twM[i];
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int k = g + j * G;
    tw[j] = k <= M / 2 ? twN[k] : mk<T>(0, 0);
  }
  const T sh = shift ? (T)*shift : (T)0;
  const NoPad P;
  const QuarterTw<T> qtw{ql, QL};
  typedef T v4 __attribute__((ext_vector_type(4)));
  cpx<T>* o = reinterpret_cast<cpx<T>*>(out);
  while (true) {
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const int k = g + j * G;
      if (k <= M / 2) {
        cpx<T> Zk, Zm;
        c2r_pre(a0[j], b0[j], tw[j], &Zk, &Zm);
        buf0[P(k)] = Zk;
        if (k > 0 && M - k != k) buf0[P(M - k)] = Zm;
        c2r_pre(a1[j], b1[j], tw[j], &Zk, &Zm);
        buf1[P(k)] = Zk;
        if (k > 0 && M - k != k) buf1[P(M - k)] = Zm;
      }
    }
    __syncthreads();
    const int yn = y + nwx;
    if (yn < yb) load(yn);  // in flight during the transforms and stores
    row_fft<T, ML, G, true>(buf0, plan, qtw, P, g);
    row_fft<T, ML, G, true>(buf1, plan, qtw, P, g);
    for (int n = g; n < M; n += G) {
      const cpx<T> a = buf0[P(n)], b = buf1[P(n)];
      const v4 val = {a.x + sh, b.x, a.y + sh, b.y};
      *reinterpret_cast<v4*>(o + tile_offset(2 * n, y, Nx)) = val;
    }
    if (yn >= yb) break;
    __syncthreads();  // every store has read the buffers
    y = yn;
  }
}

// ------------------------------------------------------------------ column kernels
// LDS of a column kernel: [twiddles W_Ny^m, m < tw_len(Ny)] [CB column buffers
// of LS elements].  FAST == false (columns too long for that layout): the
// twiddles are read from the global table.
// Global <-> LDS: thread t moves column c = t % CB, rows t/CB + i*G, so a
// wavefront touches CB contiguous elements of 64/CB rows.  Offsets fit in 32
// bits (Ny*W < 2^30 is checked on the host).

// In-place complex FFT along y of every column of an Ny x W array, CB columns
// per workgroup, times `sc`.
template <class T, int CB, int EPT, bool INV, bool FAST>
__global__ __launch_bounds__(kColThreads) void k_col_fft(cpx<T>* __restrict__ data, int Ny, int W,
                                                         T sc, Plan plan,
                                                         const cpx<T>* __restrict__ tw, int LS) {
  constexpr int G = kColThreads / CB;
  using PadT = NoPad;
  extern __shared__ __align__(16) unsigned char smem[];
  const int QL = FAST ? tw_len(Ny) : 0;
  cpx<T>* q = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* buf = q + QL;
  const int ngroups = (W + CB - 1) / CB;
  const int lb = xcd_block(ngroups);
  if (lb >= ngroups) return;
  const int t = threadIdx.x, u0 = lb * CB;
  const int c = t % CB, y0 = t / CB, u = u0 + c;
  const bool col_ok = u < W;
  const PadT P;
  for (int i = t; i < QL; i += kColThreads) q[i] = tw[i];
  for (int y = y0; y < Ny; y += G) buf[c * LS + P(y)] = col_ok ? data[y * W + u] : mk<T>(0, 0);
  __syncthreads();
  if constexpr (FAST) dev_fft<T, G, EPT, INV>(buf + (t / G) * LS, plan, QuarterTw<T>{q, QL}, P, t % G);
  else dev_fft<T, G, EPT, INV>(buf + (t / G) * LS, plan, TableTw<T>{tw}, P, t % G);
  if (col_ok)
    for (int y = y0; y < Ny; y += G) data[y * W + u] = scale(buf[c * LS + P(y)], sc);
}

// Fused per-profile column kernel.  Every profile of the reference is
// symmetric about the pixel corner its kernel is centred on (setKernel puts
// the centre at (-1/2, -1/2)), so the 2-D DFT of the wrapped kernel is
//   K[v][u] = e^{i pi (u/Nx + v/Ny)} kappa[v][u],
//   kappa[v][u] = sum_k a[k][u] cos((2k+1) pi v/Ny),   k < h = n/2,
// with a[k][u] the x-cosine sums of the kernel rows h+k and h-1-k (built by
// k_kernel_cos).  The Nyquist column u = Nx/2 of K vanishes exactly.  This
// kernel therefore evaluates K on the fly (FP64 three-term recurrence over
// the odd harmonics), multiplies by the map spectrum column and runs one
// inverse FFT along y per column and per kernel (value a, derivative ad);
// the Nyquist column of the output is written as zero.
// Columns u in [0, M), M = W-1; CB per workgroup; per-row tables csy[v] =
// (cos t, sin t, cos 2t), t = pi v/Ny, and per-column phx[u] = (cos, sin)(pi u/Nx).
template <class T, int CB, int EPT, bool FAST>
__global__ __launch_bounds__(kColThreads) void k_node_col(
    const cpx<T>* __restrict__ fmap, const double* __restrict__ av, const double* __restrict__ ad,
    int h, const double* __restrict__ csy, const double* __restrict__ phx,
    cpx<T>* __restrict__ out_v, cpx<T>* __restrict__ out_d, int Ny, int W, Plan plan,
    const cpx<T>* __restrict__ tw, int LS, int coef_lds) {
  constexpr int G = kColThreads / CB;
  constexpr int CH = 2;  // elements whose recurrences run together
  extern __shared__ __align__(16) unsigned char smem[];
  const int QL = FAST ? tw_len(Ny) : 0;
  cpx<T>* q = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* buf = q + QL;
  double* coef = reinterpret_cast<double*>(buf + (size_t)CB * LS);  // [CB][h] when coef_lds
  const int M = W - 1;
  const int ngroups = (M + CB - 1) / CB;
  const int lb = xcd_block(ngroups);
  if (lb >= ngroups) return;
  const int t = threadIdx.x, u0 = lb * CB;
  const int c = t % CB, y0 = t / CB, u = u0 + c;
  const bool col_ok = u < M;
  for (int i = t; i < QL; i += kColThreads) q[i] = tw[i];
  const double pc = col_ok ? phx[2 * u] : 0.0, ps = col_ok ? phx[2 * u + 1] : 0.0;
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    const double* a = pass ? ad : av;
    cpx<T>* dst = pass ? out_d : out_v;
    if (!dst) continue;  // uniform over the grid
    const double* ac;    // coefficients of this thread's column
    int astride;
    if (coef_lds) {
      for (int i = t; i < CB * h; i += kColThreads) {
        const int cc = i / h, k = i % h;
        coef[i] = (u0 + cc < M) ? a[(size_t)k * W + u0 + cc] : 0.0;
      }
      ac = coef + c * h;
      astride = 1;
    } else {
      ac = a + (col_ok ? u : 0);
      astride = W;
    }
    __syncthreads();
    for (int y = y0; y < Ny; y += CH * G) {
      double ct[CH], st[CH], c2[CH], Tm[CH], Tk[CH], S[CH];
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        const int ye = min(y + e * G, Ny - 1);
        ct[e] = csy[3 * ye];
        st[e] = csy[3 * ye + 1];
        c2[e] = 2.0 * csy[3 * ye + 2];
        Tm[e] = ct[e];
        Tk[e] = ct[e];
        S[e] = ac[0] * ct[e];
      }
      for (int k = 1; k < h; ++k) {
        const double ak = ac[(size_t)k * astride];
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          const double Tn = c2[e] * Tk[e] - Tm[e];
          S[e] += ak * Tn;
          Tm[e] = Tk[e];
          Tk[e] = Tn;
        }
      }
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        const int ye = y + e * G;
        if (ye < Ny) {
          // K = e^{i(phi_u + theta_v)} kappa
          const T kr = (T)(S[e] * (pc * ct[e] - ps * st[e]));
          const T ki = (T)(S[e] * (ps * ct[e] + pc * st[e]));
          buf[c * LS + ye] = col_ok ? mul(fmap[ye * W + u], mk<T>(kr, ki)) : mk<T>(0, 0);
        }
      }
    }
    __syncthreads();
    if constexpr (FAST) dev_fft<T, G, EPT, true>(buf + (t / G) * LS, plan, QuarterTw<T>{q, QL}, NoPad(), t % G);
    else dev_fft<T, G, EPT, true>(buf + (t / G) * LS, plan, TableTw<T>{tw}, NoPad(), t % G);
    if (col_ok) {
      for (int y = y0; y < Ny; y += G) dst[y * W + u] = buf[c * LS + y];
      if (u == 0)
        for (int y = y0; y < Ny; y += G) dst[y * W + M] = mk<T>(0, 0);
    }
    __syncthreads();
  }
}

// Node spectra without column FFTs.  With G the row spectrum of the map
// (unnormalised R2C of every row) and the kernel factorisation above, the
// column pass of the 2-D convolution is a short circular filter along y in
// the row-spectral domain:
//   H[y][u] = e^{i pi u/Nx}/Nx * sum_{k<h} (a[k][u]/2) (G[y-k][u] + G[y+k+1][u]),
// value taps av, derivative taps ad.  A workgroup owns kYfCols columns and
// kYfRows rows; the rows it needs (plus h above and below, circular) are
// staged in LDS, and every thread keeps kYfOut outputs and two windows of
// kYfOut + kYfK - 1 inputs in registers over each block of kYfK taps.  The
// Nyquist column (u = W-1) is zero.  Logical blocks run down each column
// strip, XCD by XCD, so the halo rows are shared in one L2.
constexpr int kYfCols = 64, kYfGroups = 16, kYfOut = 4, kYfK = 4;
constexpr int kYfRows = kYfGroups * kYfOut;

template <class T>
__global__ __launch_bounds__(kYfCols * kYfGroups) void k_node_yfilter(
    const cpx<T>* __restrict__ G, int ldg, const double* __restrict__ av, const double* __restrict__ ad,
    int h, int hp, const double* __restrict__ phx, cpx<T>* __restrict__ out_v, cpx<T>* __restrict__ out_d,
    int Nx, int Ny, int W) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* tile = reinterpret_cast<cpx<T>*>(smem);  // [kYfRows + 2 hp][kYfCols]
  cpx<T>* taps = tile + (size_t)(kYfRows + 2 * hp) * kYfCols;  // [hp][kYfCols] (value, deriv)
  const int nstrip = (W + kYfCols - 1) / kYfCols, nyb = (Ny + kYfRows - 1) / kYfRows;
  const int lb = xcd_block(nstrip * nyb);
  if (lb >= nstrip * nyb) return;
  const int u0 = (lb / nyb) * kYfCols, y0 = (lb % nyb) * kYfRows;
  const int c = threadIdx.x % kYfCols, g = threadIdx.x / kYfCols, u = u0 + c;
  const int nrow = kYfRows + 2 * hp;
  int yy = (y0 - hp + g) % Ny;  // circular row of tile row r, advanced by kYfGroups
  if (yy < 0) yy += Ny;
  const bool col_in = u < W;
  for (int r0 = g; r0 < nrow; r0 += 4 * kYfGroups) {
    // four rows per thread in flight before the LDS stores
    cpx<T> v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[e] = (col_in && r0 + e * kYfGroups < nrow) ? G[(size_t)yy * ldg + u] : mk<T>(0, 0);
      yy += kYfGroups;
      while (yy >= Ny) yy -= Ny;  // Ny may be smaller than kYfGroups
    }
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (r0 + e * kYfGroups < nrow) tile[(r0 + e * kYfGroups) * kYfCols + c] = v[e];
  }
  // the block's taps, once, in the working precision (zero beyond h)
  for (int k = g; k < hp; k += kYfGroups)
    taps[k * kYfCols + c] = (col_in && k < h)
                                ? mk<T>((T)av[(size_t)k * W + u], out_d ? (T)ad[(size_t)k * W + u] : (T)0)
                                : mk<T>(0, 0);
  __syncthreads();
  if (u >= W) return;
  const int ybase = y0 + g * kYfOut;
  const bool nyq = u == W - 1;
  cpx<T> accv[kYfOut], accd[kYfOut];
#pragma unroll
  for (int i = 0; i < kYfOut; ++i) accv[i] = accd[i] = mk<T>(0, 0);
  const cpx<T>* col = tile + c;
  for (int kb = 0; kb < hp; kb += kYfK) {
    // left[j] = G[ybase - kb - (kYfK-1) + j], right[j] = G[ybase + kb + 1 + j]
    cpx<T> L[kYfOut + kYfK - 1], R[kYfOut + kYfK - 1];
    const int lrow = g * kYfOut + hp - kb - (kYfK - 1), rrow = g * kYfOut + hp + kb + 1;
#pragma unroll
    for (int j = 0; j < kYfOut + kYfK - 1; ++j) {
      L[j] = col[(lrow + j) * kYfCols];
      R[j] = col[(rrow + j) * kYfCols];
    }
    T bv[kYfK], bd[kYfK];
#pragma unroll
    for (int kk = 0; kk < kYfK; ++kk) {
      const cpx<T> tk = taps[(kb + kk) * kYfCols + c];
      bv[kk] = tk.x;
      bd[kk] = tk.y;
    }
#pragma unroll
    for (int kk = 0; kk < kYfK; ++kk)
#pragma unroll
      for (int i = 0; i < kYfOut; ++i) {
        const cpx<T> sm = add(L[i - kk + kYfK - 1], R[i + kk]);
        accv[i] = mk<T>(accv[i].x + bv[kk] * sm.x, accv[i].y + bv[kk] * sm.y);
        accd[i] = mk<T>(accd[i].x + bd[kk] * sm.x, accd[i].y + bd[kk] * sm.y);
      }
  }
  const double sc = 0.5 / (double)Nx;
  const cpx<T> ph = mk<T>((T)(phx[2 * u] * sc), (T)(phx[2 * u + 1] * sc));
#pragma unroll
  for (int i = 0; i < kYfOut; ++i) {
    const int y = ybase + i;
    if (y < Ny) {
      out_v[(size_t)y * W + u] = nyq ? mk<T>(0, 0) : mul(accv[i], ph);
      if (out_d) out_d[(size_t)y * W + u] = nyq ? mk<T>(0, 0) : mul(accd[i], ph);
    }
  }
}

// Node-map rows from the two node spectra by one length-N complex inverse
// FFT per row (compile-time N, mlc_ctfft.hpp): with
//   Z[u] = Hv[u] + i Hd[u],  Z[N-u] = conj(Hv[u]) + i conj(Hd[u]),  0 < u < N/2
// (DC and Nyquist keep only their real parts, as a C2R does), z[x] = v[x] +
// i d[x] is the (value, d/dr) pair of pixel x.  The value gets the map mean
// back.  Rows run XCD by XCD so the four rows of a 4 x 4 tile meet in one L2.
template <class T, int N>
__global__ __launch_bounds__(CtGeom<N>::G) void k_node_rows(
    const cpx<T>* __restrict__ Hv, const cpx<T>* __restrict__ Hd, cpx<T>* __restrict__ out, int Ny,
    const double* __restrict__ shift, const cpx<T>* __restrict__ tw) {
  using Gm = CtGeom<N>;
  constexpr int G = Gm::G, W = N / 2 + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* lo = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* hi = lo + Gm::NLO;
  cpx<T>* buf = lo + Gm::NTW;
  const int y = xcd_block(Ny);
  if (y >= Ny) return;
  const int t = threadIdx.x;
  for (int i = t; i < Gm::NTW; i += G) lo[i] = tw[i];
  const cpx<T>* hv = Hv + (size_t)y * W;
  const cpx<T>* hd = Hd + (size_t)y * W;
  {
    // u = t + i G, i < 8, covers [0, N/2); all loads issued before the stores
    constexpr int NU = (N / 2) / G;
    cpx<T> a[NU], b[NU];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
      a[i] = hv[t + i * G];
      b[i] = hd[t + i * G];
    }
    cpx<T> an, bn;  // Nyquist
    if (t == 0) {
      an = hv[N / 2];
      bn = hd[N / 2];
    }
#pragma unroll
    for (int i = 0; i < NU; ++i) {
      const int u = t + i * G;
      if (u == 0) {
        buf[pad16(0)] = mk<T>(a[i].x, b[i].x);
      } else {
        buf[pad16(u)] = mk<T>(a[i].x - b[i].y, a[i].y + b[i].x);
        buf[pad16(N - u)] = mk<T>(a[i].x + b[i].y, b[i].x - a[i].y);
      }
    }
    if (t == 0) buf[pad16(N / 2)] = mk<T>(an.x, bn.x);
  }
  __syncthreads();
  ct_fft<T, N, true>(buf, lo, hi, t);
  const T sh = shift ? (T)*shift : (T)0;
  // pixel x = t + j G: consecutive lanes read consecutive LDS elements and
  // four of them fill one 32-byte row of a tile
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int x = t + j * G;
    const cpx<T> z = buf[pad16(x)];
    out[tile_offset(x, y, N)] = mk<T>(z.x + sh, z.y);
  }
}

// Node-map rows straight from the map's row spectrum: the y-filter of
// k_node_yfilter and the row transform of k_node_rows in one kernel, so the
// two node spectra never leave the chip.  A workgroup of R transform groups
// (R CtGeom<N>::G threads) owns R H consecutive rows y0 .. y0+RH-1.  First
// every thread runs the y-filter (node_filter) for its column pairs (u,
// u+1), u even, u < N/2, and all RH rows.  A pair is one 16-byte load in
// FP32 (rows of G and of the taps have even strides), which halves the
// vector-memory instructions, and each row and tap load serves RH output
// rows: the L2-to-L1 traffic of these loads, (2h + RH - 1 + h) / RH rows
// per output row, is the limit of this phase.  Each thread puts Z[u] = Hv +
// i Hd and Z[N-u] of the first R rows into the R row buffers in LDS (H = 2:
// the other R rows' sums wait in registers); then group g transforms row
// y0+g and the workgroup writes the R rows into the tiled node map; with H
// = 2 the held rows follow through the same buffers.  Taps tp are the
// (value, derivative) pairs in the working precision, hq >= h rows (zeros
// at k >= h); the sums, their order (k ascending) and the final products
// are those of k_node_yfilter and k_node_rows.  Workgroups run XCD by XCD
// over consecutive row blocks, so the halo rows of G are shared in one L2.
constexpr int kNfK = 2;

template <class T>
struct alignas(2 * sizeof(cpx<T>)) CPair {
  cpx<T> a, b;
};

// f(integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<N, I + 1>(f);
  }
}

// A raw buffer resource over a device array of fewer than 2^32 bytes: a load
// takes a per-lane byte offset and a wave-uniform byte offset, so a row
// start is scalar arithmetic and the lane's column one fixed register.
struct BufView {
  __amdgpu_buffer_rsrc_t rs;
};
__device__ __forceinline__ BufView buf_view(const void* p, uint32_t bytes) {
  // dword 3 = 0x00020000: the raw 32-bit data format of gfx9 buffer loads
  return BufView{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000)};
}
template <class T>
__device__ __forceinline__ CPair<T> buf_pair(const BufView& b, uint32_t voff, uint32_t soff) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(CPair<T>, __builtin_amdgcn_raw_buffer_load_b128(b.rs, voff, soff, 0));
  } else {
    struct W2 {
      v4u p, q;
    } w{__builtin_amdgcn_raw_buffer_load_b128(b.rs, voff, soff, 0),
        __builtin_amdgcn_raw_buffer_load_b128(b.rs, voff + 16u, soff, 0)};
    return __builtin_bit_cast(CPair<T>, w);
  }
}

// The y-filter of the fused node kernels for one column pair (u, u+1) and
// the R rows y0 .. y0+R-1:
//   accv[i][e] = sum_{k < hq} tv_k[u+e] (G[y0+i-k][u+e] + G[y0+i+k+1][u+e])
// (accd with the derivative taps), k ascending, one product-add per tap.
// The left rows G[y0+R-1-m], the right rows G[y0+1+m] (m = 0, 1, ...) and the
// tap rows sit in rings of S = R + P register slots, slot m mod S.  Tap k
// reads slots k .. k+R-1 of both rings; before its arithmetic it issues the
// loads of row m = k+R+P-1 (into the slot tap k-1 freed) and of tap k+P, so
// P taps of loads are in flight.  The loop body is unrolled over S taps, so
// every slot keeps its register and nothing is moved.  Rows and taps past
// the last ones read repeat the last ones (never used).  Row starts are
// wave-uniform byte offsets (rowb bytes per row of G, tapb per tap row);
// voff is the lane's column byte offset.  Rows wrap around Ny (Ny >= h+16).
template <class T, int R, int P>
__device__ __forceinline__ void node_filter(const BufView& gb, const BufView& tb, uint32_t voff, uint32_t rowb,
                                            uint32_t tapb, int y0, int Ny, int hq, cpx<T> (&accv)[R][2],
                                            cpx<T> (&accd)[R][2]) {
  using P2 = CPair<T>;
  constexpr int S = R + P;
  const int mmax = hq + R - 2;
  auto wrap = [&](int r) { return r < 0 ? r + Ny : (r >= Ny ? r - Ny : r); };
  auto lrow = [&](int m) { return (uint32_t)wrap(y0 + R - 1 - min(m, mmax)) * rowb; };
  auto rrow = [&](int m) { return (uint32_t)wrap(y0 + 1 + min(m, mmax)) * rowb; };
  auto trow = [&](int k) { return (uint32_t)min(k, hq - 1) * tapb; };
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int e = 0; e < 2; ++e) accv[i][e] = accd[i][e] = mk<T>(0, 0);
  P2 Ls[S], Rs[S], Ts[S];
#pragma unroll
  for (int m = 0; m < S - 1; ++m) {
    Ls[m] = buf_pair<T>(gb, voff, lrow(m));
    Rs[m] = buf_pair<T>(gb, voff, rrow(m));
  }
#pragma unroll
  for (int k = 0; k < P; ++k) Ts[k] = buf_pair<T>(tb, voff, trow(k));
  // tap k = kb + j, slot j: loads for later taps, then this tap's sums
  auto tap = [&](auto jc, int k) {
    constexpr int j = decltype(jc)::value;
    Ls[(j + S - 1) % S] = buf_pair<T>(gb, voff, lrow(k + S - 1));
    Rs[(j + S - 1) % S] = buf_pair<T>(gb, voff, rrow(k + S - 1));
    Ts[(j + P) % S] = buf_pair<T>(tb, voff, trow(k + P));
    const P2 tk = Ts[j];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const P2& l = Ls[(j + R - 1 - i) % S];
      const P2& r = Rs[(j + i) % S];
      const cpx<T> s0 = add(l.a, r.a), s1 = add(l.b, r.b);
      accv[i][0] = mk<T>(accv[i][0].x + tk.a.x * s0.x, accv[i][0].y + tk.a.x * s0.y);
      accd[i][0] = mk<T>(accd[i][0].x + tk.a.y * s0.x, accd[i][0].y + tk.a.y * s0.y);
      accv[i][1] = mk<T>(accv[i][1].x + tk.b.x * s1.x, accv[i][1].y + tk.b.x * s1.y);
      accd[i][1] = mk<T>(accd[i][1].x + tk.b.y * s1.x, accd[i][1].y + tk.b.y * s1.y);
    }
  };
  using std::integral_constant;
  // whole rounds of S taps in one basic block (no exit between taps, so the
  // compiler keeps every load where it is issued), then the last hq mod S
  int kb = 0;
  for (; kb + S <= hq; kb += S) {
    static_for<S>([&](auto jc) { tap(jc, kb + decltype(jc)::value); });
  }
  static_for<S - 1>([&](auto jc) {
    if (kb + decltype(jc)::value < hq) tap(jc, kb + decltype(jc)::value);
  });
}
// taps of loads in flight: three in FP32 (60 VGPRs of rings at R = 2), one
// in FP64 (pairs take twice the registers)
template <class T, int R, int H>
constexpr int node_filter_depth() {
  return sizeof(T) == 4 ? (H > 1 ? 1 : (R >= 4 ? 2 : 3)) : 1;
}

template <class T, int N, int R, int H>
__global__ __launch_bounds__(CtGeom<N>::G * R) void k_node_fused(
    const cpx<T>* __restrict__ Gs, int ldg, const cpx<T>* __restrict__ tp, int ldt, int hq,
    const double* __restrict__ phx, cpx<T>* __restrict__ out, int Ny, const double* __restrict__ shift,
    const cpx<T>* __restrict__ tw) {
  using Gm = CtGeom<N>;
  constexpr int G = Gm::G, NT = G * R, PPT = (N / 4) / NT, RF = R * H;
  static_assert((N / 4) % NT == 0, "column pairs split evenly over the threads");
  static_assert(H == 1 || H == 2, "one or two rounds of R rows");
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* lo = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* hi = lo + Gm::NLO;
  cpx<T>* bufs = lo + Gm::NTW;  // R buffers of Gm::LDS elements
  const int nq = (Ny + RF - 1) / RF;
  const int q = xcd_block(nq);
  if (q >= nq) return;
  const int y0 = q * RF, t = threadIdx.x;
  for (int i = t; i < Gm::NTW; i += NT) lo[i] = tw[i];
  const double sc = 0.5 / (double)N;
  const uint32_t rowb = (uint32_t)ldg * (uint32_t)sizeof(cpx<T>), tapb = (uint32_t)ldt * (uint32_t)sizeof(cpx<T>);
  const BufView gb = buf_view(Gs, (uint32_t)Ny * rowb), tb = buf_view(tp, (uint32_t)hq * tapb);
  // Z[u] = Hv + i Hd and Z[N-u] of rows r0 .. r0+R-1 from the filter sums
  // of column pair c into the R row buffers
  auto put = [&](int c, const cpx<T> (&av)[RF][2], const cpx<T> (&ad)[RF][2], int r0) {
    const int u = 2 * (t + c * NT);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int ue = u + e;
      const cpx<T> ph = mk<T>((T)(phx[2 * ue] * sc), (T)(phx[2 * ue + 1] * sc));
#pragma unroll
      for (int i = 0; i < R; ++i) {
        cpx<T>* buf = bufs + i * Gm::LDS;
        const cpx<T> a = mul(av[r0 + i][e], ph), b = mul(ad[r0 + i][e], ph);
        if (ue == 0) {
          buf[pad16(0)] = mk<T>(a.x, b.x);
        } else {
          buf[pad16(ue)] = mk<T>(a.x - b.y, a.y + b.x);
          buf[pad16(N - ue)] = mk<T>(a.x + b.y, b.x - a.y);
        }
      }
    }
  };
  // transform the R buffers (rows ya .. ya+R-1) and store them: chunk s is
  // pixels (x, x+1), x = 4 (s / 2R) + 2 (s % 2), of row ya + (s % 2R) / 2,
  // one 16-byte store; the 2R chunks of one tile are consecutive lanes and
  // fill the workgroup's R rows of that 4 x 4 tile with contiguous bytes (64
  // bytes for R = 2, the whole 128-byte tile for R = 4), instead of 32-byte
  // pieces per row.
  auto transform_store = [&](int ya) {
    if (t < R) bufs[t * Gm::LDS + pad16(N / 2)] = mk<T>(0, 0);  // Nyquist column
    __syncthreads();
    const int g = t / G, tl = t % G;
    ct_fft<T, N, true>(bufs + g * Gm::LDS, lo, hi, tl);  // ends with a barrier: all R rows done
    const T sh = shift ? (T)*shift : (T)0;
    typedef T v4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int i = 0; i < N / (2 * G); ++i) {
      const int sidx = t + i * NT, tile = sidx / (2 * R), part = sidx % (2 * R);
      const int row = part >> 1, x = 4 * tile + 2 * (part & 1), y = ya + row;
      const cpx<T>* b = bufs + row * Gm::LDS;
      const cpx<T> z0 = b[pad16(x)], z1 = b[pad16(x + 1)];
      if (y < Ny) {
        const v4 val = {z0.x + sh, z0.y, z1.x + sh, z1.y};
        node_store(reinterpret_cast<v4*>(out + tile_offset(x, y, N)), val);
      }
    }
  };
  // H = 2: rows y0+R .. y0+2R-1 are filtered with the first R (each row and
  // tap load then serves 2R output rows) and wait in registers while the
  // first R are transformed
  cpx<T> kv[H > 1 ? PPT : 1][H > 1 ? R : 1][2], kd[H > 1 ? PPT : 1][H > 1 ? R : 1][2];
#pragma unroll
  for (int c = 0; c < PPT; ++c) {
    const int u = 2 * (t + c * NT);
    cpx<T> accv[RF][2], accd[RF][2];
    node_filter<T, RF, node_filter_depth<T, RF, H>()>(gb, tb, (uint32_t)u * (uint32_t)sizeof(cpx<T>), rowb, tapb,
                                                   y0, Ny, hq, accv, accd);
    put(c, accv, accd, 0);
    if constexpr (H > 1) {
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          kv[c][i][e] = accv[R + i][e];
          kd[c][i][e] = accd[R + i][e];
        }
    }
  }
  transform_store(y0);
  if constexpr (H > 1) {
    __syncthreads();  // every store has read the buffers
#pragma unroll
    for (int c = 0; c < PPT; ++c) {
      cpx<T> av[RF][2], ad[RF][2];
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          av[R + i][e] = kv[c][i][e];
          ad[R + i][e] = kd[c][i][e];
        }
      put(c, av, ad, R);
    }
    transform_store(y0 + R);
  }
}

// Node-map rows of a width N that is not a power of two (N = 2^a 3^b 5^c,
// a compile-time constant): the y-filter of k_node_fused and one complex
// inverse FFT of length N per row with the compile-time plan of fixed_fft
// (radices 8, 4, 2, 5, 3; quarter-wave twiddle table in LDS), G threads per
// row (a multiple of 64 with 16 G >= N), R rows per workgroup.  Column
// pairs and store chunks that the thread count does not divide evenly are
// predicated.  The filter's sums and their order are those of k_node_fused.
struct Pad16 {
  MLC_HD int operator()(int i) const { return pad16(i); }
};

template <class T, int N, int R, int G>
__global__ __launch_bounds__(G * R) void k_node_fused_mix(
    const cpx<T>* __restrict__ Gs, int ldg, const cpx<T>* __restrict__ tp, int ldt, int hq,
    const double* __restrict__ phx, cpx<T>* __restrict__ out, int Ny, const double* __restrict__ shift,
    const cpx<T>* __restrict__ twN) {
  constexpr int NT = G * R, QL = N / 4, LDSN = N + N / 16;
  static_assert(N % 4 == 0 && G % 64 == 0 && 16 * G >= N, "row geometry");
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* qt = reinterpret_cast<cpx<T>*>(smem);  // W_N^m, m < N/4
  cpx<T>* bufs = qt + QL;                         // R buffers of LDSN elements
  const int nq = (Ny + R - 1) / R;
  const int q = xcd_block(nq);
  if (q >= nq) return;  // uniform per workgroup
  const int y0 = q * R, t = threadIdx.x;
  for (int i = t; i < QL; i += NT) qt[i] = twN[i];
  const double sc = 0.5 / (double)N;
  const uint32_t rowb = (uint32_t)ldg * (uint32_t)sizeof(cpx<T>), tapb = (uint32_t)ldt * (uint32_t)sizeof(cpx<T>);
  const BufView gb = buf_view(Gs, (uint32_t)Ny * rowb), tb = buf_view(tp, (uint32_t)hq * tapb);
  for (int cp = t; cp < N / 4; cp += NT) {
    const int u = 2 * cp;
    cpx<T> ph[2];
#pragma unroll
    for (int e = 0; e < 2; ++e)
      ph[e] = mk<T>((T)(phx[2 * (u + e)] * sc), (T)(phx[2 * (u + e) + 1] * sc));
    cpx<T> accv[R][2], accd[R][2];
    node_filter<T, R, node_filter_depth<T, R, 1>()>(gb, tb, (uint32_t)u * (uint32_t)sizeof(cpx<T>), rowb, tapb,
                                                y0, Ny, hq, accv, accd);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      cpx<T>* buf = bufs + i * LDSN;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int ue = u + e;
        const cpx<T> a = mul(accv[i][e], ph[e]), b = mul(accd[i][e], ph[e]);
        if (ue == 0) {
          buf[pad16(0)] = mk<T>(a.x, b.x);
        } else {
          buf[pad16(ue)] = mk<T>(a.x - b.y, a.y + b.x);
          buf[pad16(N - ue)] = mk<T>(a.x + b.y, b.x - a.y);
        }
      }
    }
  }
  if (t < R) bufs[t * LDSN + pad16(N / 2)] = mk<T>(0, 0);  // Nyquist column
  __syncthreads();
  const int g = t / G, tl = t % G;
  fixed_fft<T, N, G, 16, true, 0, 1>(bufs + g * LDSN, QuarterTw<T>{qt, QL}, Pad16{}, tl);  // ends with a barrier
  const T sh = shift ? (T)*shift : (T)0;
  typedef T v4 __attribute__((ext_vector_type(4)));
  for (int sidx = t; sidx < (N / 2) * R; sidx += NT) {
    const int tile = sidx / (2 * R), part = sidx % (2 * R);
    const int row = part >> 1, x = 4 * tile + 2 * (part & 1), y = y0 + row;
    const cpx<T>* b = bufs + row * LDSN;
    const cpx<T> z0 = b[pad16(x)], z1 = b[pad16(x + 1)];
    if (y < Ny) {
      const v4 val = {z0.x + sh, z0.y, z1.x + sh, z1.y};
      node_store(reinterpret_cast<v4*>(out + tile_offset(x, y, N)), val);
    }
  }
}

// The map's row spectrum with the compile-time FFT of M = Nx/2 points (one
// row per workgroup): the row minus its own FP64 mean, packed as z[n] =
// x[2n] + i x[2n+1], one forward DFT_M, then the split into the half
// spectrum X[0..M] (r2c_post).  The FP64 row sum goes to rowsum[y];
// k_row_dc_fix sets the DC terms afterwards.
template <class T, int M>
__global__ __launch_bounds__(CtGeom<M>::G) void k_map_rows(const double* __restrict__ in,
                                                         cpx<T>* __restrict__ out, int ldo,
                                                         double* __restrict__ rowsum,
                                                         const cpx<T>* __restrict__ tw,
                                                         const cpx<T>* __restrict__ twN) {
  using Gm = CtGeom<M>;
  constexpr int G = Gm::G, NU = M / G;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[(G + 63) / 64];
  cpx<T>* lo = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* hi = lo + Gm::NLO;
  cpx<T>* buf = lo + Gm::NTW;
  const int y = blockIdx.x, t = threadIdx.x;
  for (int i = t; i < Gm::NTW; i += G) lo[i] = tw[i];
  const cpx<double>* row = reinterpret_cast<const cpx<double>*>(in + (size_t)y * 2 * M);
  cpx<double> v[NU];
#pragma unroll
  for (int i = 0; i < NU; ++i) v[i] = row[t + i * G];
  double a = 0.0;
#pragma unroll
  for (int i = 0; i < NU; ++i) a += v[i].x + v[i].y;
  // workgroups of fewer than 64 threads reduce over their own lanes only
  for (int o = (G < 64 ? G : 64) / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if ((t & 63) == 0) red[t >> 6] = a;
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int w = 0; w < (G + 63) / 64; ++w) tot += red[w];
  if (t == 0) rowsum[y] = tot;
  const double mean = tot / (2.0 * M);
#pragma unroll
  for (int i = 0; i < NU; ++i)
    buf[pad16(t + i * G)] = mk<T>((T)(v[i].x - mean), (T)(v[i].y - mean));
  __syncthreads();
  ct_fft<T, M, false>(buf, lo, hi, t);
  cpx<T>* o = out + (size_t)y * ldo;
  for (int k = t; k <= M / 2; k += G) {
    cpx<T> Xk, Xm;
    r2c_post(buf[pad16(k)], buf[pad16(k == 0 ? 0 : M - k)], twN[k], &Xk, &Xm);
    o[k] = Xk;
    if (M - k != k) o[M - k] = Xm;
  }
}

// ------------------------------------------------------------------ profiles
// One workgroup evaluates the n x n profile and its d/drhalf, normalises both
// (K = f/S, dK = (df - K dS)/S) in FP64.
__device__ __forceinline__ void profile_block(const ProfileParams& P, double* __restrict__ prof,
                                              double* __restrict__ dprof) {
  __shared__ double red[2][1024];
  const int n = P.n, t = threadIdx.x;
  double s = 0.0, ds = 0.0;
  for (int idx = t; idx < n * n; idx += 1024) {
    int i = idx / n, j = idx % n;
    double x = profile_coord(j, n, P.pix), y = profile_coord(i, n, P.pix);
    double f, df;
    profile_value(P, hypot(x, y), &f, &df);
    prof[idx] = f;
    dprof[idx] = df;
    s += f;
    ds += df;
  }
  red[0][t] = s;
  red[1][t] = ds;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] += red[0][t + w];
      red[1][t] += red[1][t + w];
    }
    __syncthreads();
  }
  const double S = red[0][0], dS = red[1][0];
  for (int idx = t; idx < n * n; idx += 1024) {
    double K = prof[idx] / S;
    dprof[idx] = (dprof[idx] - K * dS) / S;
    prof[idx] = K;
  }
}

__global__ __launch_bounds__(1024) void k_profile(ProfileParams P, double* __restrict__ prof,
                                                  double* __restrict__ dprof) {
  profile_block(P, prof, dprof);
}

// One node map of a call (the nodes of one map, or of several maps of the
// batched call): its profile grid and d/drhalf (n x n each), its cosine
// coefficients (h x W each, FP64) and its taps in the working precision
// (hq x ldt complex elements), for a map of width Nx with W = Nx/2 + 1 bins
// and the cosine table cosx of that width.
struct NodeJob {
  ProfileParams P;
  double *prof, *dprof, *av, *ad;
  void* tp;
  const double* cosx;
  int Nx, W, ldt;
};

// One workgroup per job.
__global__ __launch_bounds__(1024) void k_profile_many(const NodeJob* __restrict__ jobs) {
  const NodeJob& J = jobs[blockIdx.x];
  profile_block(J.P, J.prof, J.dprof);
}

// Cosine coefficients of the symmetric kernel (see k_node_col):
//   a[k][u] = rho[h+k][u] + rho[h-1-k][u],
//   rho[s][u] = sum_c P[s][c] cos(pi u (2(c-h)+1)/Nx),
// from the table cosx[m] = cos(pi m/Nx), m < 2 Nx.  FP64.  grid (ceil(W/256), h).
template <class T>
__device__ __forceinline__ void kernel_cos_at(const double* __restrict__ prof,
                                              const double* __restrict__ dprof, int n, int Nx, int W,
                                              const double* __restrict__ cosx, double* __restrict__ av,
                                              double* __restrict__ ad, cpx<T>* __restrict__ tp, int ldt,
                                              int u, int k) {
  const int h = n >> 1;
  const int s1 = h + k, s2 = h - 1 - k, twoN = 2 * Nx;
  double sa = 0.0, sd = 0.0;
  // angle index u (2(cc-h)+1) mod 2Nx, advanced by 2u per column
  int m = (int)(((long long)u * (1 - 2 * h) % twoN + twoN) % twoN);
  const int step = (int)((2LL * u) % twoN);
  for (int cc = 0; cc < n; ++cc) {
    const double w = cosx[m];
    m += step;
    if (m >= twoN) m -= twoN;
    sa += (prof[s1 * n + cc] + prof[s2 * n + cc]) * w;
    sd += (dprof[s1 * n + cc] + dprof[s2 * n + cc]) * w;
  }
  av[(size_t)k * W + u] = sa;
  if (ad) ad[(size_t)k * W + u] = sd;
  // (value, derivative) taps in the working precision for k_node_fused
  if (tp) tp[(size_t)k * ldt + u] = mk<T>((T)sa, (T)sd);
}

template <class T>
__global__ __launch_bounds__(256) void k_kernel_cos(const double* __restrict__ prof,
                                                    const double* __restrict__ dprof, int n, int Nx,
                                                    int W, const double* __restrict__ cosx,
                                                    double* __restrict__ av, double* __restrict__ ad,
                                                    cpx<T>* __restrict__ tp, int ldt) {
  const int u = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
  if (u >= W) return;
  if (k >= (n >> 1)) {  // the zero taps up to a multiple of kNfK
    if (tp) tp[(size_t)k * ldt + u] = mk<T>(0, 0);
    return;
  }
  kernel_cos_at<T>(prof, dprof, n, Nx, W, cosx, av, ad, tp, ldt, u, k);
}

// grid (ceil(max W / 256), max h rounded up to kNfK, jobs).
template <class T>
__global__ __launch_bounds__(256) void k_kernel_cos_many(const NodeJob* __restrict__ jobs) {
  const NodeJob& J = jobs[blockIdx.z];
  const int n = J.P.n, u = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y, h = n >> 1;
  if (u >= J.W) return;
  cpx<T>* tp = static_cast<cpx<T>*>(J.tp);
  if (k >= h) {  // the zero taps up to a multiple of kNfK
    if (k < (h + kNfK - 1) / kNfK * kNfK) tp[(size_t)k * J.ldt + u] = mk<T>(0, 0);
    return;
  }
  kernel_cos_at<T>(J.prof, J.dprof, n, J.Nx, J.W, J.cosx, J.av, J.ad, tp, J.ldt, u, k);
}

// Full Ny x Nx wrapped kernel (Kernel::setKernel), for the plain path.
template <class T>
__global__ void k_make_kernel(const double* __restrict__ prof, int n, int Nx, int Ny,
                              T* __restrict__ ker) {
  const int h = n >> 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)Nx * Ny;
       i += (int64_t)gridDim.x * blockDim.x) {
    int y = (int)(i / Nx), x = (int)(i % Nx);
    int s = wrap_row(y, h, Ny), c = wrap_row(x, h, Nx);
    ker[i] = (s < n && c < n) ? (T)prof[(size_t)s * n + c] : (T)0;
  }
}

// kernel_multiplyFFTs (gpu_functions.cu:571-577) without the hard-coded grid;
// the 1/(Nx Ny) factor is already part of the map spectrum.
template <class T>
__global__ void k_cmul(const cpx<T>* __restrict__ fmap, cpx<T>* __restrict__ fk, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    fk[i] = mul(fmap[i], fk[i]);
}

// (value, deriv) of the plain path into the tiled node map.
template <class T>
__global__ void k_interleave(const T* __restrict__ a, const T* __restrict__ b, cpx<T>* __restrict__ o,
                             int Nx, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    o[tile_offset((int)(i % Nx), (int)(i / Nx), Nx)] = mk<T>(a[i], b[i]);
}

// ------------------------------------------------------------------ host side
template <class T>
const cpx<T>* twiddles(mlc_ctx* ctx, int L) {
  auto key = std::make_pair(L, (int)sizeof(T));
  auto it = ctx->twiddles.find(key);
  if (it != ctx->twiddles.end()) return (const cpx<T>*)it->second;
  std::vector<cpx<T>> h(L);
  for (int m = 0; m < L; ++m) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)L;
    h[m] = mk<T>((T)std::cos(a), (T)std::sin(a));
  }
  void* d = nullptr;
  hip_check(hipMalloc(&d, sizeof(cpx<T>) * L), "hipMalloc(twiddles)");
  ctx->twiddles[key] = d;
  hip_check(hipMemcpy(d, h.data(), sizeof(cpx<T>) * L, hipMemcpyHostToDevice), "hipMemcpy(twiddles)");
  return (const cpx<T>*)d;
}

// FP64 host-built tables for the cosine-sum kernel spectrum (cached per
// context): kind 1: cos(pi m/N), m < 2N;  kind 2: (cos t, sin t, cos 2t),
// t = pi v/N, v < N;  kind 3: (cos, sin)(pi u/N), u <= N/2.
const double* cos_table(mlc_ctx* ctx, int kind, int N) {
  auto key = std::make_pair(N, 1000 + kind);
  auto it = ctx->twiddles.find(key);
  if (it != ctx->twiddles.end()) return (const double*)it->second;
  const long double pi = 3.14159265358979323846264338327950288L;
  std::vector<double> h;
  if (kind == 1) {
    for (int m = 0; m < 2 * N; ++m) h.push_back((double)std::cos(pi * m / N));
  } else if (kind == 2) {
    for (int v = 0; v < N; ++v) {
      long double t = pi * v / N;
      h.push_back((double)std::cos(t));
      h.push_back((double)std::sin(t));
      h.push_back((double)std::cos(2 * t));
    }
  } else {
    for (int u = 0; u <= N / 2; ++u) {
      long double t = pi * u / N;
      h.push_back((double)std::cos(t));
      h.push_back((double)std::sin(t));
    }
  }
  void* d = nullptr;
  hip_check(hipMalloc(&d, sizeof(double) * h.size()), "hipMalloc(cos table)");
  ctx->twiddles[key] = d;
  hip_check(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice),
            "hipMemcpy(cos table)");
  return (const double*)d;
}

Plan plan_for(int L, const char* what, int max_radix = 16) {
  Plan p;
  if (L < 1 || !make_plan(L, &p, max_radix))
    fail(MLC_ERR_UNSUPPORTED, std::string(what) + ": length " + std::to_string(L) +
                                  " has a prime factor other than 2, 3, 5");
  return p;
}

template <class K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    hip_check(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes),
              "hipFuncSetAttribute(LDS)");
}

// Row launches: G threads with G*16 >= M (16 elements per thread at most).
template <class F>
void with_row_group(int M, const char* what, F&& f) {
  using std::integral_constant;
  if (M <= 64 * 16) f(integral_constant<int, 64>{});
  else if (M <= 256 * 16) f(integral_constant<int, 256>{});
  else if (M <= 1024 * 16) f(integral_constant<int, 1024>{});
  else fail(MLC_ERR_UNSUPPORTED, std::string(what) + ": row too long");
}

template <class T>
size_t row_lds(int M, int nbuf) {
  return sizeof(cpx<T>) * ((size_t)tw_len(M) + (size_t)nbuf * M);
}

// Calls f(integral_constant<int, M>) for the half lengths with a
// compile-time row kernel (the widths 1000, 1280, 1536, 2000, 2560, 3072,
// 5000 and 10000 of the workloads and of the reference's maps), else
// f(integral_constant<int, 0>) (runtime plan).
template <class F>
void with_fixed_len(int M, F&& f) {
  using std::integral_constant;
  switch (M) {
    case 500: f(integral_constant<int, 500>{}); break;
    case 640: f(integral_constant<int, 640>{}); break;
    case 768: f(integral_constant<int, 768>{}); break;
    case 1000: f(integral_constant<int, 1000>{}); break;
    case 1280: f(integral_constant<int, 1280>{}); break;
    case 1536: f(integral_constant<int, 1536>{}); break;
    case 2500: f(integral_constant<int, 2500>{}); break;
    case 5000: f(integral_constant<int, 5000>{}); break;
    default: f(integral_constant<int, 0>{}); break;
  }
}

template <class T, class TIN>
void row_r2c(mlc_ctx* ctx, const TIN* in, cpx<T>* out, int Nx, int Ny, const double* shift,
             double* rowsum = nullptr, int ldo = 0) {
  if (ldo == 0) ldo = Nx / 2 + 1;
  const int M = Nx / 2;
  Plan p = plan_for(M, "row FFT", 8);
  const cpx<T>* twM = twiddles<T>(ctx, M);
  const cpx<T>* twN = twiddles<T>(ctx, Nx);
  const size_t lds = row_lds<T>(M, 1);
  if (lds > kLdsBytes) fail(MLC_ERR_UNSUPPORTED, "row FFT: row does not fit in LDS");
  with_row_group(M, "row FFT", [&](auto gc) {
    constexpr int G = decltype(gc)::value;
    with_fixed_len(M, [&](auto lc) {
      constexpr int ML = decltype(lc)::value;
      if constexpr (ML == 0 || ML <= G * 16) {
        auto k = k_row_r2c<T, TIN, G, ML>;
        allow_lds(k, lds);
        MLC_LAUNCH(ctx, "fft_row_r2c", k, dim3(Ny), dim3(G), lds, in, out, ldo, Nx, shift, rowsum, p, twM,
                   twN);
      }
    });
  });
}

template <class T, int MODE>
void row_c2r(mlc_ctx* ctx, const cpx<T>* in0, const cpx<T>* in1, T* out, int Nx, int Ny,
             const double* shift) {
  const int M = Nx / 2;
  Plan p = plan_for(M, "row FFT", 8);
  const cpx<T>* twM = twiddles<T>(ctx, M);
  const cpx<T>* twN = twiddles<T>(ctx, Nx);
  if constexpr (MODE == 2) {
    if (row_lds<T>(M, 2) > kLdsBytes) {  // one member of the pairs per pass
      row_c2r<T, 3>(ctx, in0, nullptr, out, Nx, Ny, shift);
      row_c2r<T, 4>(ctx, in1, nullptr, out, Nx, Ny, nullptr);
      return;
    }
  }
  const size_t lds = row_lds<T>(M, MODE == 2 ? 2 : 1);
  if (lds > kLdsBytes) fail(MLC_ERR_UNSUPPORTED, "row FFT: row does not fit in LDS");
  if constexpr (MODE == 2 && sizeof(T) == 4) {
    if (M == 5000) {
      // 10000-wide node rows: the twiddles from the global table, so that
      // two workgroups (2 x 80 KB of row buffers) share a CU
      const size_t lds2 = sizeof(cpx<T>) * ((size_t)tw_len(M) + 2 * (size_t)M);
      auto k = k_row_c2r_node_pf<T, 1024, 5000>;
      allow_lds(k, lds2);
      // one workgroup per CU, a multiple of 8 (one share per XCD)
      const int grid = (int)std::min<int64_t>(((int64_t)Ny + 7) / 8 * 8, cdiv(fill_grid(ctx, 1), 8) * 8);
      MLC_LAUNCH(ctx, "fft_row_c2r_node", k, dim3(grid), dim3(1024), lds2, in0, in1, out, Ny, shift, p, twM,
                 twN);
      return;
    }
  }
  with_row_group(M, "row FFT", [&](auto gc) {
    constexpr int G = decltype(gc)::value;
    with_fixed_len(M, [&](auto lc) {
      constexpr int ML = decltype(lc)::value;
      if constexpr (ML == 0 || ML <= G * 16) {
        auto k = k_row_c2r<T, G, MODE, ML>;
        allow_lds(k, lds);
        const int grid = MODE >= 2 ? ((Ny + 7) / 8) * 8 : Ny;
        MLC_LAUNCH(ctx, MODE >= 2 ? "fft_row_c2r_node" : "fft_row_c2r", k, dim3(grid), dim3(G), lds, in0,
                   in1, out, Nx, Ny, shift, p, twM, twN);
      }
    });
  });
}

// Column launch geometry: CB adjacent columns per 1024-thread workgroup, at
// most EPT elements per thread; FAST = padded buffers + LDS twiddles.
struct ColGeom {
  int CB, EPT, LS;
  bool fast;
  size_t lds;
};

// Geometries with a kernel instantiation, in order of preference (widest CB
// first, then fewest elements per thread); the unpadded, global-twiddle
// layout is the fallback for columns too long for the fast one.
constexpr int kColCases[][3] = {{16, 8, 1}, {8, 8, 1}, {8, 16, 1}, {4, 8, 1}, {4, 16, 1},
                                {2, 8, 1},  {2, 16, 1}, {2, 24, 1}, {1, 8, 1}, {1, 16, 1},
                                {1, 24, 1}, {1, 16, 0}, {1, 24, 0}};

template <class T>
ColGeom col_geom(int Ny) {
  for (auto& cs : kColCases) {
    const int CB = cs[0], EPT = cs[1];
    const bool fast = cs[2] != 0;
    const int LS = Ny + 1;  // odd stride: neighbouring columns start on different banks
    const size_t lds = sizeof(cpx<T>) * ((size_t)LS * CB + (fast ? tw_len(Ny) : 0));
    if (lds <= kLdsBytes && Ny * CB <= kColThreads * EPT) return {CB, EPT, LS, fast, lds};
  }
  fail(MLC_ERR_UNSUPPORTED, "column FFT: column of " + std::to_string(Ny) + " does not fit in LDS");
}

template <class F>
void with_col(const ColGeom& g, F&& f) {
  using std::integral_constant;
#define MLC_COL_CASE(cb, ept, fst)                                                         \
  if (g.CB == cb && g.EPT == ept && g.fast == (fst != 0)) {                                \
    f(integral_constant<int, cb>{}, integral_constant<int, ept>{}, integral_constant<bool, fst != 0>{}); \
    return;                                                                                \
  }
  MLC_COL_CASE(16, 8, 1) MLC_COL_CASE(8, 8, 1) MLC_COL_CASE(8, 16, 1) MLC_COL_CASE(4, 8, 1)
  MLC_COL_CASE(4, 16, 1) MLC_COL_CASE(2, 8, 1) MLC_COL_CASE(2, 16, 1) MLC_COL_CASE(2, 24, 1)
  MLC_COL_CASE(1, 8, 1) MLC_COL_CASE(1, 16, 1) MLC_COL_CASE(1, 24, 1) MLC_COL_CASE(1, 16, 0)
  MLC_COL_CASE(1, 24, 0)
#undef MLC_COL_CASE
  fail(MLC_ERR_UNSUPPORTED, "column FFT: no kernel for this geometry");
}

template <class T, bool INV>
void col_fft(mlc_ctx* ctx, cpx<T>* data, int Ny, int W, T scale) {
  Plan p = plan_for(Ny, "column FFT", 8);
  const cpx<T>* tw = twiddles<T>(ctx, Ny);
  ColGeom g = col_geom<T>(Ny);
  with_col(g, [&](auto cbc, auto eptc, auto fc) {
    constexpr int CB = decltype(cbc)::value, EPT = decltype(eptc)::value;
    constexpr bool FAST = decltype(fc)::value;
    int ngroups = (W + CB - 1) / CB;
    auto k = k_col_fft<T, CB, EPT, INV, FAST>;
    allow_lds(k, g.lds);
    MLC_LAUNCH(ctx, INV ? "fft_col_inv" : "fft_col_fwd", k, dim3(((ngroups + 7) / 8) * 8),
               dim3(kColThreads), g.lds, data, Ny, W, scale, p, tw, g.LS);
  });
}

// The node kernel covers the M = W-1 columns below the (zero) Nyquist column.
template <class T>
void node_col(mlc_ctx* ctx, const cpx<T>* fmap, const double* av, const double* ad, int h,
              cpx<T>* ov, cpx<T>* od, int Nx, int Ny, int W) {
  Plan p = plan_for(Ny, "column FFT", 8);
  const cpx<T>* tw = twiddles<T>(ctx, Ny);
  const double* csy = cos_table(ctx, 2, Ny);
  const double* phx = cos_table(ctx, 3, Nx);
  ColGeom g = col_geom<T>(Ny);
  const size_t coef = sizeof(double) * (size_t)g.CB * h;
  const int coef_lds = g.lds + coef <= kLdsBytes;
  const size_t lds = g.lds + (coef_lds ? coef : 0);
  with_col(g, [&](auto cbc, auto eptc, auto fc) {
    constexpr int CB = decltype(cbc)::value, EPT = decltype(eptc)::value;
    constexpr bool FAST = decltype(fc)::value;
    int ngroups = (W - 1 + CB - 1) / CB;
    auto k = k_node_col<T, CB, EPT, FAST>;
    allow_lds(k, lds);
    MLC_LAUNCH(ctx, "conv_node_col", k, dim3(((ngroups + 7) / 8) * 8), dim3(kColThreads), lds,
               fmap, av, ad, h, csy, phx, ov, od, Ny, W, p, tw, g.LS, coef_lds);
  });
}

// Twiddle tables of mlc_ctfft.hpp for length N: lo[l] = w^l, l < 2^LB,
// then hi[h] = w^(h 2^LB), then t2[16 r + k] = w^(k r N/256), k, r < 16,
// with w = exp(+2 pi i/N) for the inverse transform and exp(-2 pi i/N) for
// the forward one (the stages use the table as given), each entry rounded
// once from long double; cached per context and direction.
template <class T, int N, bool INV>
const cpx<T>* ct_twiddles(mlc_ctx* ctx) {
  using Gm = CtGeom<N>;
  auto key = std::make_pair(N, (INV ? 3000 : 4000) + (int)sizeof(T));
  auto it = ctx->twiddles.find(key);
  if (it != ctx->twiddles.end()) return (const cpx<T>*)it->second;
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double sg = INV ? 1.0L : -1.0L;
  std::vector<cpx<T>> h;
  for (int l = 0; l < Gm::NLO; ++l)
    h.push_back(mk<T>((T)std::cos(sg * 2 * pi * l / N), (T)std::sin(sg * 2 * pi * l / N)));
  for (int k = 0; k < Gm::NHI; ++k) {
    const long double a = sg * 2 * pi * (long double)k * Gm::NLO / N;
    h.push_back(mk<T>((T)std::cos(a), (T)std::sin(a)));
  }
  for (int r = 0; r < 16; ++r)  // t2: the second stage's w^(k r N/256)
    for (int k = 0; k < 16; ++k) {
      const long double a = sg * 2 * pi * (long double)(k * r) / 256;
      h.push_back(mk<T>((T)std::cos(a), (T)std::sin(a)));
    }
  void* d = nullptr;
  hip_check(hipMalloc(&d, sizeof(cpx<T>) * h.size()), "hipMalloc(twiddles)");
  ctx->twiddles[key] = d;
  hip_check(hipMemcpy(d, h.data(), sizeof(cpx<T>) * h.size(), hipMemcpyHostToDevice),
            "hipMemcpy(twiddles)");
  return (const cpx<T>*)d;
}

// Node maps from the node spectra with the compile-time row FFT; false for
// widths it does not cover (not a power of two in [256, 16384], or rows
// that do not fit in LDS).
// k_map_rows for Nx = 2M with M a power of two from 256 to 8192; false for
// other widths (the runtime-plan row kernel takes them).
template <class T>
bool map_rows(mlc_ctx* ctx, const double* in, cpx<T>* out, int ldo, double* rowsum, int Nx, int Ny) {
  bool done = false;
  auto go = [&](auto mc) {
    constexpr int M = decltype(mc)::value;
    using Gm = CtGeom<M>;
    const size_t lds = sizeof(cpx<T>) * (size_t)(Gm::NTW + Gm::LDS);
    if (lds > kLdsBytes - 1024) return;
    auto k = k_map_rows<T, M>;
    allow_lds(k, lds);
    MLC_LAUNCH(ctx, "fft_row_r2c", k, dim3(Ny), dim3(Gm::G), lds, in, out, ldo, rowsum, ct_twiddles<T, M, false>(ctx),
               twiddles<T>(ctx, Nx));
    done = true;
  };
  using std::integral_constant;
  switch (Nx / 2) {
    case 256: go(integral_constant<int, 256>{}); break;
    case 512: go(integral_constant<int, 512>{}); break;
    case 1024: go(integral_constant<int, 1024>{}); break;
    case 2048: go(integral_constant<int, 2048>{}); break;
    case 4096: go(integral_constant<int, 4096>{}); break;
    case 8192: go(integral_constant<int, 8192>{}); break;
    default: break;
  }
  return done;
}

template <class T>
bool node_rows(mlc_ctx* ctx, const cpx<T>* Hv, const cpx<T>* Hd, T* out, int Nx, int Ny,
               const double* shift) {
  bool done = false;
  auto go = [&](auto nc) {
    constexpr int N = decltype(nc)::value;
    using Gm = CtGeom<N>;
    const size_t lds = sizeof(cpx<T>) * (size_t)(Gm::NTW + Gm::LDS);
    if (lds > kLdsBytes) return;
    auto k = k_node_rows<T, N>;
    allow_lds(k, lds);
    MLC_LAUNCH(ctx, "fft_row_c2r_node", k, dim3(((Ny + 7) / 8) * 8), dim3(Gm::G), lds, Hv, Hd,
               (cpx<T>*)out, Ny, shift, ct_twiddles<T, N, true>(ctx));
    done = true;
  };
  using std::integral_constant;
  switch (Nx) {
    case 256: go(integral_constant<int, 256>{}); break;
    case 512: go(integral_constant<int, 512>{}); break;
    case 1024: go(integral_constant<int, 1024>{}); break;
    case 2048: go(integral_constant<int, 2048>{}); break;
    case 4096: go(integral_constant<int, 4096>{}); break;
    case 8192: go(integral_constant<int, 8192>{}); break;
    case 16384: go(integral_constant<int, 16384>{}); break;
    default: break;
  }
  return done;
}

// Node map from the row spectrum by k_node_fused; false where the separate
// y-filter and row kernels must run: widths outside 256 .. 16384, a row that
// does not fit in LDS, fewer than h + 16 rows, or taps that the y-filter
// would not stage either (so both paths take the column FFTs for the same
// profiles).  Rows per workgroup: 4 up to 1024 columns, else 2 where two
// row buffers fit, else 1.  At 4096 FP32 columns two row buffers (70 KB of
// LDS, two workgroups per CU, one filtering while the other transforms)
// beat four (140 KB, one workgroup per CU, no overlap); at 2048 and 4096
// FP32 columns each workgroup filters four rows through its two buffers
// (H = 2), so each row and tap load serves four output rows.  ldg and ldt
// (row strides of Gs and tp) must be even, and the row spectrum below 4 GiB.
// With launch = false nothing is launched: the return value says whether the
// call would launch k_node_fused.
template <class T>
bool node_fused(mlc_ctx* ctx, const cpx<T>* Gs, int ldg, const cpx<T>* tp, int ldt, int h, T* out, int Nx,
                int Ny, const double* shift, bool launch = true) {
  const int hp = (h + kYfK - 1) / kYfK * kYfK;
  if (sizeof(cpx<T>) * (size_t)(kYfRows + 3 * hp) * kYfCols > kLdsBytes) return false;
  if (Ny < h + 16) return false;
  // the filter's buffer loads address G with 32-bit byte offsets
  if ((size_t)Ny * ldg * sizeof(cpx<T>) >= (size_t(1) << 32)) return false;
  bool done = false;
  auto go = [&](auto nc) {
    constexpr int N = decltype(nc)::value;
    using Gm = CtGeom<N>;
    auto lds_for = [](int r) { return sizeof(cpx<T>) * (size_t)(Gm::NTW + r * Gm::LDS); };
    auto run = [&](auto rc, auto hc) {
      constexpr int R = decltype(rc)::value, H = decltype(hc)::value;
      done = true;
      if (!launch) return;
      const size_t lds = lds_for(R);
      auto k = k_node_fused<T, N, R, H>;
      allow_lds(k, lds);
      const int nq = (Ny + R * H - 1) / (R * H);
      MLC_LAUNCH(ctx, "conv_node_fused", k, dim3(((nq + 7) / 8) * 8), dim3(Gm::G * R), lds, Gs, ldg, tp, ldt,
                 h, cos_table(ctx, 3, Nx), (cpx<T>*)out, Ny, shift, ct_twiddles<T, N, true>(ctx));
    };
    using std::integral_constant;
    if constexpr (N <= 1024) {
      run(integral_constant<int, 4>{}, integral_constant<int, 1>{});
    } else if constexpr (sizeof(T) == 4 && (N == 2048 || N == 4096)) {
      run(integral_constant<int, 2>{}, integral_constant<int, 2>{});
    } else {
      if (lds_for(2) <= kLdsBytes) run(integral_constant<int, 2>{}, integral_constant<int, 1>{});
      else if (lds_for(1) <= kLdsBytes) run(integral_constant<int, 1>{}, integral_constant<int, 1>{});
    }
  };
  // widths that are not powers of two: G threads per row, R rows per
  // workgroup (halved in FP64, whose rows take twice the LDS)
  auto mix = [&](auto nc, auto gc, auto rc) {
    constexpr int N = decltype(nc)::value, G = decltype(gc)::value;
    constexpr int R = sizeof(T) == 4 ? decltype(rc)::value : (decltype(rc)::value > 1 ? decltype(rc)::value / 2 : 1);
    const size_t lds = sizeof(cpx<T>) * ((size_t)N / 4 + (size_t)R * (N + N / 16));
    if (lds > kLdsBytes) return;
    done = true;
    if (!launch) return;
    auto k = k_node_fused_mix<T, N, R, G>;
    allow_lds(k, lds);
    const int nq = (Ny + R - 1) / R;
    MLC_LAUNCH(ctx, "conv_node_fused", k, dim3(((nq + 7) / 8) * 8), dim3(G * R), lds, Gs, ldg, tp, ldt, h,
               cos_table(ctx, 3, Nx), (cpx<T>*)out, Ny, shift, twiddles<T>(ctx, N));
  };
  using std::integral_constant;
  switch (Nx) {
    case 256: go(integral_constant<int, 256>{}); break;
    case 512: go(integral_constant<int, 512>{}); break;
    case 1024: go(integral_constant<int, 1024>{}); break;
    case 2048: go(integral_constant<int, 2048>{}); break;
    case 4096: go(integral_constant<int, 4096>{}); break;
    case 8192: go(integral_constant<int, 8192>{}); break;
    case 16384: go(integral_constant<int, 16384>{}); break;
    case 1000: mix(integral_constant<int, 1000>{}, integral_constant<int, 64>{}, integral_constant<int, 4>{}); break;
    case 1280: mix(integral_constant<int, 1280>{}, integral_constant<int, 128>{}, integral_constant<int, 4>{}); break;
    case 1536: mix(integral_constant<int, 1536>{}, integral_constant<int, 128>{}, integral_constant<int, 4>{}); break;
    case 2000: mix(integral_constant<int, 2000>{}, integral_constant<int, 128>{}, integral_constant<int, 4>{}); break;
    case 2560: mix(integral_constant<int, 2560>{}, integral_constant<int, 192>{}, integral_constant<int, 2>{}); break;
    case 3072: mix(integral_constant<int, 3072>{}, integral_constant<int, 192>{}, integral_constant<int, 2>{}); break;
    default: break;
  }
  return done;
}

// y-filter node spectra from the row spectrum; false when the staged rows
// do not fit in LDS (the caller then takes the column-FFT path).
template <class T>
bool node_yfilter(mlc_ctx* ctx, const cpx<T>* rows, int ldg, const double* av, const double* ad, int h,
                  cpx<T>* ov, cpx<T>* od, int Nx, int Ny, int W) {
  const int hp = (h + kYfK - 1) / kYfK * kYfK;
  const size_t lds = sizeof(cpx<T>) * (size_t)(kYfRows + 3 * hp) * kYfCols;
  if (lds > kLdsBytes) return false;
  const int nb = ((W + kYfCols - 1) / kYfCols) * ((Ny + kYfRows - 1) / kYfRows);
  auto k = k_node_yfilter<T>;
  allow_lds(k, lds);
  MLC_LAUNCH(ctx, "conv_node_yfilter", k, dim3(((nb + 7) / 8) * 8), dim3(kYfCols * kYfGroups), lds,
             rows, ldg, av, ad, h, hp, cos_table(ctx, 3, Nx), ov, od, Nx, Ny, W);
  return true;
}

}  // namespace

namespace mlc {

// Whether the per-map work of a batched call runs side by side on the spread
// streams: for maps of up to 4096^2 (small launches that leave most of the
// chip idle; a step over 64 maps of 1024^2 to 4096^2 gains 17%, with each
// map given to the stream with the fewest pixels so far, largest first).  The work
// of larger maps fills the chip by itself, and four 10000^2 maps side by
// side were 3% slower than in turn.  A map listed twice keeps the calls in
// turn (its buffers would be shared by two streams).
inline bool spread_maps(const std::vector<mlc_map*>& maps) {
  for (size_t i = 0; i < maps.size(); ++i) {
    if ((int64_t)maps[i]->Nx * maps[i]->Ny > (int64_t)4096 * 4096) return false;
    for (size_t j = 0; j < i; ++j)
      if (maps[j] == maps[i]) return false;
  }
  return maps.size() > 1;
}

template <class T>
void fft_maps_transform(const std::vector<mlc_map*>& maps, const double* const* d_maps) {
  mlc_ctx* ctx = maps[0]->ctx;
  const int nm = (int)maps.size();
  Workspace ws(ctx);
  std::vector<int> ir(nm, -1);
  for (int i = 0; i < nm; ++i)
    if (maps[i]->subtract_mean) ir[i] = ws.add(sizeof(double) * maps[i]->Ny);
  ws.commit();
  // every row transform (maps of up to 4096^2 side by side on the spread
  // streams), then the mean and DC step of all mean-subtracting maps in one
  // launch (the same sums as fft_map_transform's)
  auto xform = [&](int i) {
    mlc_map* m = maps[i];
    if (m->subtract_mean) {
      double* rowsum = ws.get<double>(ir[i]);
      if (!map_rows<T>(ctx, d_maps[i], (cpx<T>*)m->rows, m->ldr, rowsum, m->Nx, m->Ny))
        row_r2c<T, double>(ctx, d_maps[i], (cpx<T>*)m->rows, m->Nx, m->Ny, nullptr, rowsum, m->ldr);
    } else {
      row_r2c<T, double>(ctx, d_maps[i], (cpx<T>*)m->rows, m->Nx, m->Ny, nullptr, nullptr, m->ldr);
    }
  };
  if (spread_maps(maps)) {
    std::vector<double> w(nm);
    for (int i = 0; i < nm; ++i) w[i] = (double)maps[i]->Nx * maps[i]->Ny;
    run_spread(ctx, nm, xform, &w);
  } else {
    for (int i = 0; i < nm; ++i) xform(i);
  }
  std::vector<DcJob> dc;
  for (int i = 0; i < nm; ++i) {
    mlc_map* m = maps[i];
    if (m->subtract_mean) {
      double* rowsum = ws.get<double>(ir[i]);
      DcJob J;
      std::memset(&J, 0, sizeof J);  // (padding too: see node_tables)
      J.rowsum = rowsum;
      J.out = m->rows;
      J.mean = m->mean;
      J.Nx = m->Nx;
      J.Ny = m->Ny;
      J.ldo = m->ldr;
      dc.push_back(J);
    }
    m->full_ready = false;
  }
  if (!dc.empty()) {
    const DcJob* dj = static_cast<const DcJob*>(const_table(ctx, kTableDcJobs, dc.data(), sizeof(DcJob) * dc.size()));
    MLC_LAUNCH(ctx, "map_dc", k_row_dc_fix_many<T>, dim3((int)dc.size()), dim3(1024), 0, dj);
  }
}

template <class T>
void fft_map_transform(mlc_map* m, const double* d_map) {
  mlc_ctx* ctx = m->ctx;
  if (m->subtract_mean) {
    // each row is transformed minus its own mean (the sums come out of the
    // same pass over the map); the map mean and the DC terms follow from the
    // row sums
    Workspace ws(ctx);
    int ir = ws.add(sizeof(double) * m->Ny);
    ws.commit();
    double* rowsum = ws.get<double>(ir);
    if (!map_rows<T>(ctx, d_map, (cpx<T>*)m->rows, m->ldr, rowsum, m->Nx, m->Ny))
      row_r2c<T, double>(ctx, d_map, (cpx<T>*)m->rows, m->Nx, m->Ny, nullptr, rowsum, m->ldr);
    MLC_LAUNCH(ctx, "map_dc", k_row_dc_fix<T>, dim3(1), dim3(1024), 0, rowsum, m->Nx, m->Ny,
               (cpx<T>*)m->rows, m->ldr, m->mean);
  } else {
    // m->mean was zeroed once when the map was created and stays zero, so an
    // update issues kernels only (no memset node under graph capture)
    row_r2c<T, double>(ctx, d_map, (cpx<T>*)m->rows, m->Nx, m->Ny, nullptr, nullptr, m->ldr);
  }
  m->full_ready = false;
}

// The full 2-D spectrum / (Nx Ny) in m->fmap, from the row spectrum, on
// first use after a map update (the plain path and the column-FFT node path).
template <class T>
void ensure_full(mlc_map* m) {
  if (m->full_ready) return;
  const size_t e = sizeof(cpx<T>);
  hip_check(hipMemcpy2DAsync(m->fmap, e * m->W, m->rows, e * m->ldr, e * m->W, m->Ny, hipMemcpyDeviceToDevice,
                             m->ctx->stream),
            "hipMemcpy2DAsync(row spectrum)");
  col_fft<T, false>(m->ctx, (cpx<T>*)m->fmap, m->Ny, m->W,
                    (T)(1.0 / ((double)m->Nx * (double)m->Ny)));
  m->full_ready = true;
}

template <class T>
void fft_convolve_fused(mlc_map* m, const ProfileParams& P, void* d_value, void* d_node) {
  mlc_ctx* ctx = m->ctx;
  const int n = P.n, W = m->W;
  const bool want_d = d_node != nullptr;
  const int h = n / 2;
  Workspace ws(ctx);
  int ip = ws.add(sizeof(double) * n * n), idp = ws.add(sizeof(double) * n * n);
  int ia = ws.add(sizeof(double) * h * W), iad = ws.add(sizeof(double) * h * W);
  const int ldt = W + (W & 1), hq = (h + kNfK - 1) / kNfK * kNfK;
  int itp = ws.add(sizeof(cpx<T>) * hq * ldt);
  ws.commit();
  double *prof = ws.get<double>(ip), *dprof = ws.get<double>(idp);
  double *av = ws.get<double>(ia), *ad = ws.get<double>(iad);
  cpx<T>* tp = ws.get<cpx<T>>(itp);
  MLC_LAUNCH(ctx, "profile", k_profile, dim3(1), dim3(1024), 0, P, prof, dprof);
  MLC_LAUNCH(ctx, "conv_kernel_cos", k_kernel_cos<T>, dim3(cdiv(W, 256), want_d ? hq : h), dim3(256), 0, prof,
             dprof, n,
             m->Nx, W, cos_table(ctx, 1, m->Nx), av, want_d ? ad : nullptr, want_d ? tp : nullptr, ldt);
  if (d_node && !d_value &&
      node_fused<T>(ctx, (const cpx<T>*)m->rows, m->ldr, tp, ldt, h, (T*)d_node, m->Nx, m->Ny, m->mean))
    return;
  cpx<T>* G0 = (cpx<T>*)m->g0;
  cpx<T>* G1 = (cpx<T>*)m->g1;
  if (!node_yfilter<T>(ctx, (const cpx<T>*)m->rows, m->ldr, av, ad, h, G0, want_d ? G1 : nullptr, m->Nx, m->Ny,
                       W)) {
    ensure_full<T>(m);
    node_col<T>(ctx, (const cpx<T>*)m->fmap, av, ad, h, G0, want_d ? G1 : nullptr, m->Nx, m->Ny, W);
  }
  if (d_node && !node_rows<T>(ctx, G0, G1, (T*)d_node, m->Nx, m->Ny, m->mean))
    row_c2r<T, 2>(ctx, G0, G1, (T*)d_node, m->Nx, m->Ny, m->mean);
  if (d_value) row_c2r<T, 1>(ctx, G0, nullptr, (T*)d_value, m->Nx, m->Ny, m->mean);
}

// Node maps of several profiles: the profiles and their cosine
// coefficients in one launch each, then the y-filter and row transforms of
// each node (the same kernels and arithmetic as fft_convolve_fused), in
// order or in the order a driver asks for them.  The profile tables live in
// the map's own buffer, so a driver may run work that uses the context
// workspace between the nodes.
// Profile grids, cosine coefficients and taps of the nodes Ps of map m, in
// m's node-table storage (one slice each): one NodeJob per node, appended to
// jobs; hmax and wmax grow to the largest tap rows (a multiple of kNfK) and
// the widest bin count.
template <class T>
void node_tables(mlc_map* m, const std::vector<ProfileParams>& Ps, std::vector<NodeJob>& jobs, int& hmax,
                 int& wmax) {
  mlc_ctx* ctx = m->ctx;
  const int cnt = (int)Ps.size(), W = m->W, ldt = W + (W & 1);
  Workspace ws(ctx, &m->ntab, &m->ntab_bytes);
  std::vector<int> sl(5 * (size_t)cnt);
  for (int k = 0; k < cnt; ++k) {
    const int n = Ps[k].n, h = n / 2, hq = (h + kNfK - 1) / kNfK * kNfK;
    sl[5 * k] = ws.add(sizeof(double) * n * n);
    sl[5 * k + 1] = ws.add(sizeof(double) * n * n);
    sl[5 * k + 2] = ws.add(sizeof(double) * std::max<int64_t>((int64_t)h * W, 1));
    sl[5 * k + 3] = ws.add(sizeof(double) * std::max<int64_t>((int64_t)h * W, 1));
    sl[5 * k + 4] = ws.add(sizeof(cpx<T>) * std::max<int64_t>((int64_t)hq * ldt, 1));
    hmax = std::max(hmax, hq);
  }
  wmax = std::max(wmax, W);
  ws.commit();
  const double* cosx = cos_table(ctx, 1, m->Nx);
  for (int k = 0; k < cnt; ++k) {
    // zeroed first: the device table is rewritten only when its bytes,
    // padding included, change
    NodeJob J;
    std::memset(&J, 0, sizeof J);
    J.P = Ps[k];
    J.prof = ws.get<double>(sl[5 * k]);
    J.dprof = ws.get<double>(sl[5 * k + 1]);
    J.av = ws.get<double>(sl[5 * k + 2]);
    J.ad = ws.get<double>(sl[5 * k + 3]);
    J.tp = ws.get<cpx<T>>(sl[5 * k + 4]);
    J.cosx = cosx;
    J.Nx = m->Nx;
    J.W = W;
    J.ldt = ldt;
    jobs.push_back(J);
  }
}

// The profiles and taps of every job in the device table dj (njobs jobs).
template <class T>
void launch_node_tables(mlc_ctx* ctx, const NodeJob* dj, int njobs, int hmax, int wmax) {
  MLC_LAUNCH(ctx, "profile", k_profile_many, dim3(njobs), dim3(1024), 0, dj);
  MLC_LAUNCH(ctx, "conv_kernel_cos", k_kernel_cos_many<T>, dim3(cdiv(wmax, 256), hmax, njobs), dim3(256), 0,
             dj);
}

// Node map J of map m into dst: the fused kernel, else the separate y-filter
// (or the column-FFT path) and row kernels.
template <class T>
void convolve_node(mlc_map* m, const NodeJob& J, void* dst) {
  mlc_ctx* ctx = m->ctx;
  const int h = J.P.n / 2, W = m->W;
  if (node_fused<T>(ctx, (const cpx<T>*)m->rows, m->ldr, static_cast<const cpx<T>*>(J.tp), J.ldt, h, (T*)dst,
                    m->Nx, m->Ny, m->mean))
    return;
  cpx<T>* G0 = (cpx<T>*)m->g0;
  cpx<T>* G1 = (cpx<T>*)m->g1;
  if (!node_yfilter<T>(ctx, (const cpx<T>*)m->rows, m->ldr, J.av, J.ad, h, G0, G1, m->Nx, m->Ny, W)) {
    ensure_full<T>(m);
    node_col<T>(ctx, (const cpx<T>*)m->fmap, J.av, J.ad, h, G0, G1, m->Nx, m->Ny, W);
  }
  if (!node_rows<T>(ctx, G0, G1, (T*)dst, m->Nx, m->Ny, m->mean))
    row_c2r<T, 2>(ctx, G0, G1, (T*)dst, m->Nx, m->Ny, m->mean);
}

template <class T>
void fft_convolve_nodes(mlc_map* m, const std::vector<ProfileParams>& Ps, void* const* d_nodes,
                        const NodeDriver* drive) {
  mlc_ctx* ctx = m->ctx;
  const int cnt = (int)Ps.size();
  if (cnt == 0) {
    if (drive) (*drive)([](int) {}, true);
    return;
  }
  std::vector<NodeJob> jobs;
  int hmax = 1, wmax = 1;
  node_tables<T>(m, Ps, jobs, hmax, wmax);
  const NodeJob* dj =
      static_cast<const NodeJob*>(const_table_in(ctx, m->jobs, jobs.data(), sizeof(NodeJob) * cnt));
  launch_node_tables<T>(ctx, dj, cnt, hmax, wmax);
  auto conv = [&](int k) {
    MLC_REQUIRE(k >= 0 && k < cnt, "convolve_nodes: node index out of range");
    convolve_node<T>(m, jobs[k], d_nodes[k]);
  };
  if (drive) {
    // every node by k_node_fused alone: the nodes share no scratch buffer
    bool independent = true;
    for (const NodeJob& J : jobs)
      independent = independent && node_fused<T>(ctx, (const cpx<T>*)m->rows, m->ldr, static_cast<const cpx<T>*>(J.tp),
                                                  J.ldt, J.P.n / 2, (T*)d_nodes[0], m->Nx, m->Ny, m->mean, false);
    (*drive)(conv, independent);
  } else {
    for (int k = 0; k < cnt; ++k) conv(k);
  }
}

template <class T>
void fft_maps_convolve_nodes(const std::vector<mlc_map*>& maps, const std::vector<std::vector<ProfileParams>>& Ps,
                             void* const* d_nodes) {
  mlc_ctx* ctx = maps[0]->ctx;
  std::vector<NodeJob> jobs;
  std::vector<int> first;  // first job of each map
  int hmax = 1, wmax = 1;
  for (size_t i = 0; i < maps.size(); ++i) {
    first.push_back((int)jobs.size());
    if (!Ps[i].empty()) node_tables<T>(maps[i], Ps[i], jobs, hmax, wmax);
  }
  if (jobs.empty()) return;
  const NodeJob* dj =
      static_cast<const NodeJob*>(const_table(ctx, kTableNodeJobs, jobs.data(), sizeof(NodeJob) * jobs.size()));
  launch_node_tables<T>(ctx, dj, (int)jobs.size(), hmax, wmax);
  // maps of up to 4096^2 side by side on the spread streams; a map's nodes
  // in order on one stream (they share its y-filter buffers)
  auto conv = [&](int i) {
    for (size_t k = 0; k < Ps[i].size(); ++k) convolve_node<T>(maps[i], jobs[first[i] + k], d_nodes[first[i] + k]);
  };
  if (spread_maps(maps)) {
    std::vector<double> w(maps.size());
    for (size_t i = 0; i < maps.size(); ++i) w[i] = (double)maps[i]->Nx * maps[i]->Ny * Ps[i].size();
    run_spread(ctx, (int)maps.size(), conv, &w);
  } else {
    for (size_t i = 0; i < maps.size(); ++i) conv((int)i);
  }
}

template <class T>
void fft_convolve_plain(mlc_map* m, const ProfileParams& P, void* d_value, void* d_node) {
  // The reference's structure (gpu_functions.cu:184-208): full-size kernel,
  // forward FFT, multiply, inverse FFT, each stage followed by a sync.
  mlc_ctx* ctx = m->ctx;
  const int n = P.n, W = m->W, Nx = m->Nx, Ny = m->Ny;
  const int64_t npix = (int64_t)Nx * Ny, nspec = (int64_t)Ny * W;
  Workspace ws(ctx);
  int ip = ws.add(sizeof(double) * n * n), idp = ws.add(sizeof(double) * n * n);
  int ik = ws.add(sizeof(T) * npix), ifk = ws.add(sizeof(cpx<T>) * nspec);
  int iv = ws.add(sizeof(T) * npix), id = ws.add(sizeof(T) * npix);
  ws.commit();
  double *prof = ws.get<double>(ip), *dprof = ws.get<double>(idp);
  T* ker = ws.get<T>(ik);
  cpx<T>* fk = ws.get<cpx<T>>(ifk);
  T* outv = d_value ? (T*)d_value : ws.get<T>(iv);
  T* outd = ws.get<T>(id);
  ensure_full<T>(m);
  MLC_LAUNCH(ctx, "profile", k_profile, dim3(1), dim3(1024), 0, P, prof, dprof);
  sync(ctx);
  const int grid = std::min<int64_t>(cdiv(npix, 256), fill_grid(ctx, 16));
  for (int pass = 0; pass < (d_node ? 2 : 1); ++pass) {
    T* dst = pass ? outd : outv;
    MLC_LAUNCH(ctx, "plain_make_kernel", k_make_kernel<T>, dim3(grid), dim3(256), 0,
               pass ? dprof : prof, n, Nx, Ny, ker);
    sync(ctx);
    row_r2c<T, T>(ctx, ker, fk, Nx, Ny, nullptr);
    col_fft<T, false>(ctx, fk, Ny, W, (T)1);
    sync(ctx);
    MLC_LAUNCH(ctx, "plain_cmul", k_cmul<T>, dim3(grid), dim3(256), 0, (const cpx<T>*)m->fmap, fk,
               nspec);
    sync(ctx);
    col_fft<T, true>(ctx, fk, Ny, W, (T)1);
    if (pass == 0) row_c2r<T, 1>(ctx, fk, nullptr, dst, Nx, Ny, m->mean);
    else row_c2r<T, 0>(ctx, fk, nullptr, dst, Nx, Ny, nullptr);
    sync(ctx);
  }
  if (d_node) {
    MLC_LAUNCH(ctx, "plain_interleave", k_interleave<T>, dim3(grid), dim3(256), 0, outv, outd,
               (cpx<T>*)d_node, Nx, npix);
    sync(ctx);
  }
}

template <class T>
void fft_r2c_2d(mlc_ctx* ctx, const T* in, cpx<T>* out, int Nx, int Ny) {
  row_r2c<T, T>(ctx, in, out, Nx, Ny, nullptr);
  col_fft<T, false>(ctx, out, Ny, Nx / 2 + 1, (T)1);
}

// inverse of a half spectrum held in `work` (overwritten) into `out`
template <class T>
void fft_c2r_2d(mlc_ctx* ctx, cpx<T>* work, T* out, int Nx, int Ny) {
  col_fft<T, true>(ctx, work, Ny, Nx / 2 + 1, (T)1);
  row_c2r<T, 0>(ctx, work, nullptr, out, Nx, Ny, nullptr);
}

#define MLC_FFT_INSTANTIATE(T)                                                               \
  template void fft_map_transform<T>(mlc_map*, const double*);                               \
  template void fft_maps_transform<T>(const std::vector<mlc_map*>&, const double* const*);   \
  template void fft_maps_convolve_nodes<T>(const std::vector<mlc_map*>&,                     \
                                           const std::vector<std::vector<ProfileParams>>&,   \
                                           void* const*);                                    \
  template void fft_convolve_fused<T>(mlc_map*, const ProfileParams&, void*, void*);         \
  template void fft_convolve_nodes<T>(mlc_map*, const std::vector<ProfileParams>&, void* const*,    \
                                      const NodeDriver*);                                    \
  template void fft_convolve_plain<T>(mlc_map*, const ProfileParams&, void*, void*);         \
  template void fft_r2c_2d<T>(mlc_ctx*, const T*, cpx<T>*, int, int);                        \
  template void fft_c2r_2d<T>(mlc_ctx*, cpx<T>*, T*, int, int);

}  // namespace mlc
