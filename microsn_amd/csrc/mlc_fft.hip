// mlc_fft.hip -- the library's own 2-D FFT and the fused map x profile
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
#include <algorithm>
#include <cmath>
#include <vector>

#include "mlc_fft_core.hpp"
#include "mlc_internal.hpp"
#include "mlc_profile.hpp"

using namespace mlc;

namespace {

constexpr int kColThreads = 1024;
constexpr size_t kLdsBytes = 160 * 1024;

// ------------------------------------------------------------------ device FFT
template <class T, int R, int G, int EPT, bool INV>
__device__ __forceinline__ void dev_stage(cpx<T>* buf, int L, int Ns, const cpx<T>* __restrict__ tw,
                                          int g) {
  constexpr int BMAX = (EPT + R - 1) / R;
  cpx<T> v[BMAX][R];
  const int nb = L / R;
#pragma unroll
  for (int i = 0; i < BMAX; ++i) {
    int j = g + i * G;
    if (j < nb) stage_load<T, R, INV>(buf, L, Ns, j, tw, v[i]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < BMAX; ++i) {
    int j = g + i * G;
    if (j < nb) stage_store<T, R>(buf, Ns, j, v[i]);
  }
  __syncthreads();
}

// All stages of `plan` on buf[0..plan.L) by the G threads of one group
// (g = index in the group).  Every thread of the workgroup must call this
// (the stages synchronise the whole workgroup); groups run the same plan.
template <class T, int G, int EPT, bool INV>
__device__ void dev_fft(cpx<T>* buf, const Plan& plan, const cpx<T>* __restrict__ tw, int g) {
  int Ns = 1;
  for (int s = 0; s < plan.n; ++s) {
    const int R = plan.r[s];
    if (R == 8) dev_stage<T, 8, G, EPT, INV>(buf, plan.L, Ns, tw, g);
    else if (R == 4) dev_stage<T, 4, G, EPT, INV>(buf, plan.L, Ns, tw, g);
    else if (R == 2) dev_stage<T, 2, G, EPT, INV>(buf, plan.L, Ns, tw, g);
    else if (R == 5) dev_stage<T, 5, G, EPT, INV>(buf, plan.L, Ns, tw, g);
    else dev_stage<T, 3, G, EPT, INV>(buf, plan.L, Ns, tw, g);
    Ns *= R;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs; this numbering gives
// each XCD a contiguous range of logical blocks.  The grid is launched with
// 8*ceil(nb/8) workgroups; ids >= nb return at once.
__device__ __forceinline__ int xcd_block(int nb) {
  const int per = (nb + 7) >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ int wrap_row(int y, int h, int N) {
  int s = y + h;
  return s >= N ? s - N : s;
}

// ------------------------------------------------------------------ row kernels
// Forward real FFT of one row per workgroup.  in: Ny x Nx of TIN; out: Ny x W.
// With shift != NULL the FP64 value *shift is subtracted before rounding to T.
template <class T, class TIN, int G, int EPT>
__global__ __launch_bounds__(G) void k_row_r2c(const TIN* __restrict__ in, cpx<T>* __restrict__ out,
                                               int Nx, const double* __restrict__ shift, Plan plan,
                                               const cpx<T>* __restrict__ twM,
                                               const cpx<T>* __restrict__ twN) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);
  const int y = blockIdx.x, g = threadIdx.x, M = Nx >> 1, W = M + 1;
  const cpx<TIN>* row = reinterpret_cast<const cpx<TIN>*>(in + (size_t)y * Nx);
  const double sh = shift ? *shift : 0.0;
  for (int n = g; n < M; n += G) {
    cpx<TIN> p = row[n];
    buf[n] = mk<T>((T)((double)p.x - sh), (T)((double)p.y - sh));
  }
  __syncthreads();
  dev_fft<T, G, EPT, false>(buf, plan, twM, g);
  cpx<T>* o = out + (size_t)y * W;
  for (int k = g; k <= (M >> 1); k += G) {
    cpx<T> Xk, Xm;
    r2c_post(buf[k == 0 ? 0 : k], buf[k == 0 ? 0 : M - k], twN[k], &Xk, &Xm);
    o[k] = Xk;
    if (M - k != k) o[M - k] = Xm;
  }
}

// Loads the folded inverse input of one row into buf (see c2r_pre).
template <class T, int G>
__device__ __forceinline__ void c2r_fill(cpx<T>* buf, const cpx<T>* __restrict__ X, int M,
                                         const cpx<T>* __restrict__ twN, int g) {
  for (int k = g; k <= (M >> 1); k += G) {
    cpx<T> Zk, Zm;
    c2r_pre(X[k], X[M - k], twN[k], &Zk, &Zm);
    buf[k] = Zk;
    if (k > 0 && M - k != k) buf[M - k] = Zm;
  }
}

// Inverse real FFT of one row per workgroup (unnormalised).
// MODE 0: out is T[Ny][Nx] = x;  MODE 1: out is T[Ny][Nx] = x + *shift;
// MODE 2: out is (T,T)[Ny][Nx] = (x0 + *shift, x1) from the spectra in0, in1.
template <class T, int G, int EPT, int MODE>
__global__ __launch_bounds__(G) void k_row_c2r(const cpx<T>* __restrict__ in0,
                                               const cpx<T>* __restrict__ in1, T* __restrict__ out,
                                               int Nx, const double* __restrict__ shift, Plan plan,
                                               const cpx<T>* __restrict__ twM,
                                               const cpx<T>* __restrict__ twN) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf0 = reinterpret_cast<cpx<T>*>(smem);
  const int y = blockIdx.x, g = threadIdx.x, M = Nx >> 1, W = M + 1;
  c2r_fill<T, G>(buf0, in0 + (size_t)y * W, M, twN, g);
  __syncthreads();
  dev_fft<T, G, EPT, true>(buf0, plan, twM, g);
  const T sh = (MODE != 0 && shift) ? (T)*shift : (T)0;
  if constexpr (MODE == 2) {
    cpx<T>* buf1 = buf0 + M;
    c2r_fill<T, G>(buf1, in1 + (size_t)y * W, M, twN, g);
    __syncthreads();
    dev_fft<T, G, EPT, true>(buf1, plan, twM, g);
    // (value, deriv) pairs of samples 2n and 2n+1: 4 contiguous T
    cpx<T>* o = reinterpret_cast<cpx<T>*>(out) + (size_t)y * Nx;
    for (int n = g; n < M; n += G) {
      cpx<T> a = buf0[n], b = buf1[n];
      o[2 * n] = mk<T>(a.x + sh, b.x);
      o[2 * n + 1] = mk<T>(a.y + sh, b.y);
    }
  } else {
    cpx<T>* o = reinterpret_cast<cpx<T>*>(out + (size_t)y * Nx);
    for (int n = g; n < M; n += G) {
      cpx<T> a = buf0[n];
      o[n] = mk<T>(a.x + sh, a.y + sh);
    }
  }
}

// ------------------------------------------------------------------ column kernels
// In-place complex FFT along y of every column of an Ny x W array, CB columns
// per workgroup, times `scale`.
template <class T, int CB, int EPT, bool INV>
__global__ __launch_bounds__(kColThreads) void k_col_fft(cpx<T>* __restrict__ data, int Ny, int W,
                                                         T sc, Plan plan,
                                                         const cpx<T>* __restrict__ tw, int LS) {
  constexpr int G = kColThreads / CB;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);
  const int ngroups = (W + CB - 1) / CB;
  const int lb = xcd_block(ngroups);
  if (lb >= ngroups) return;
  // global <-> LDS: thread t moves column c = t % CB, rows t/CB + i*G, so a
  // wavefront touches CB contiguous elements of 64/CB rows.
  const int u0 = lb * CB, t = threadIdx.x;
  const int c = t % CB, y0 = t / CB, u = u0 + c;
  const bool col_ok = u < W;
  for (int y = y0; y < Ny; y += G) buf[c * LS + y] = col_ok ? data[y * W + u] : mk<T>(0, 0);
  __syncthreads();
  dev_fft<T, G, EPT, INV>(buf + (t / G) * LS, plan, tw, t % G);
  if (col_ok)
    for (int y = y0; y < Ny; y += G) data[y * W + u] = scale(buf[c * LS + y], sc);
}

// Fused per-profile column kernel.  For each wanted kernel (value: kr, d/dr:
// kd; both n x W rows of the x-transformed wrapped kernel, row s holding map
// row y = s - n/2 mod Ny) and each of the CB columns:
//   buf = kernel column -> FFT_y -> * Fmap column -> inverse FFT_y -> out.
// With EPT <= 8 the map spectrum column is read once into registers; with
// more elements per thread it is re-read per pass (L2-resident) to avoid
// register spills at 1024 threads per workgroup.
template <class T, int CB, int EPT>
__global__ __launch_bounds__(kColThreads) void k_node_col(
    const cpx<T>* __restrict__ fmap, const cpx<T>* __restrict__ kr, const cpx<T>* __restrict__ kd,
    int n, cpx<T>* __restrict__ out_v, cpx<T>* __restrict__ out_d, int Ny, int W, Plan plan,
    const cpx<T>* __restrict__ tw, int LS) {
  constexpr int G = kColThreads / CB;
  constexpr bool KEEP = EPT <= 8;
  constexpr int NF = KEEP ? EPT : 1;
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem);
  const int ngroups = (W + CB - 1) / CB;
  const int lb = xcd_block(ngroups);
  if (lb >= ngroups) return;
  const int u0 = lb * CB, t = threadIdx.x, h = n >> 1;
  // element i of this thread: column c (fixed, kColThreads % CB == 0), row
  // y = y0 + i*G.  Offsets fit in 32 bits (Ny*W < 2^31 is checked on the host).
  const int c = t % CB, y0 = t / CB, u = u0 + c;
  const bool col_ok = u < W;
  cpx<T> f[NF];
  if constexpr (KEEP) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      int y = y0 + i * G;
      f[i] = (y < Ny && col_ok) ? fmap[y * W + u] : mk<T>(0, 0);
    }
  }
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    const cpx<T>* src = pass ? kd : kr;
    cpx<T>* dst = pass ? out_d : out_v;
    if (!dst) continue;  // uniform over the grid
    for (int y = y0; y < Ny; y += G) {
      int s = wrap_row(y, h, Ny);
      buf[c * LS + y] = (s < n && col_ok) ? src[s * W + u] : mk<T>(0, 0);
    }
    __syncthreads();
    dev_fft<T, G, EPT, false>(buf + (t / G) * LS, plan, tw, t % G);
    if constexpr (KEEP) {
#pragma unroll
      for (int i = 0; i < EPT; ++i) {
        int y = y0 + i * G;
        if (y < Ny) buf[c * LS + y] = mul(buf[c * LS + y], f[i]);
      }
    } else {
      for (int y = y0; y < Ny; y += G) {
        cpx<T> fv = col_ok ? fmap[y * W + u] : mk<T>(0, 0);
        buf[c * LS + y] = mul(buf[c * LS + y], fv);
      }
    }
    __syncthreads();
    dev_fft<T, G, EPT, true>(buf + (t / G) * LS, plan, tw, t % G);
    if (col_ok)
      for (int y = y0; y < Ny; y += G) dst[y * W + u] = buf[c * LS + y];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ profiles
// One workgroup evaluates the n x n profile and its d/drhalf, normalises both
// (K = f/S, dK = (df - K dS)/S) in FP64.
__global__ __launch_bounds__(1024) void k_profile(ProfileParams P, double* __restrict__ prof,
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

// x-transform of the wrapped kernel rows: R[s][u] = sum_c P[s][c] W_Nx^(u (c-n/2)),
// accumulated in FP64 and rounded to T.  grid (ceil(W/256), n).
template <class T>
__global__ __launch_bounds__(256) void k_kernel_rows(const double* __restrict__ prof,
                                                     const double* __restrict__ dprof, int n, int Nx,
                                                     int W, const cpx<double>* __restrict__ twN64,
                                                     cpx<T>* __restrict__ R, cpx<T>* __restrict__ dR) {
  const int u = blockIdx.x * 256 + threadIdx.x, s = blockIdx.y, h = n >> 1;
  if (u >= W) return;
  double ar = 0, ai = 0, br = 0, bi = 0;
  const double* p = prof + (size_t)s * n;
  const double* dp = dprof + (size_t)s * n;
  for (int c = 0; c < n; ++c) {
    long long m = ((long long)u * (c - h)) % Nx;
    if (m < 0) m += Nx;
    cpx<double> w = twN64[m];
    ar += p[c] * w.x;
    ai += p[c] * w.y;
    br += dp[c] * w.x;
    bi += dp[c] * w.y;
  }
  R[(size_t)s * W + u] = mk<T>((T)ar, (T)ai);
  if (dR) dR[(size_t)s * W + u] = mk<T>((T)br, (T)bi);
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

template <class T>
__global__ void k_interleave(const T* __restrict__ a, const T* __restrict__ b, cpx<T>* __restrict__ o,
                             int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    o[i] = mk<T>(a[i], b[i]);
}

// FP64 sum of a map: per-workgroup partials, then one workgroup adds them in
// order (deterministic).  out[0] = sum / npix.
__global__ __launch_bounds__(256) void k_sum_partial(const double* __restrict__ x, int64_t n,
                                                     double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_final(const double* __restrict__ part, int np,
                                                   double inv_n, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * inv_n;
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

Plan plan_for(int L, const char* what) {
  Plan p;
  if (L < 1 || !make_plan(L, &p))
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

// Row launches: G threads and at most EPT elements per thread (G*EPT >= M).
template <class F>
void with_row_group(int M, const char* what, F&& f) {
  using std::integral_constant;
  if (M <= 64 * 8) f(integral_constant<int, 64>{}, integral_constant<int, 8>{});
  else if (M <= 256 * 8) f(integral_constant<int, 256>{}, integral_constant<int, 8>{});
  else if (M <= 1024 * 8) f(integral_constant<int, 1024>{}, integral_constant<int, 8>{});
  else if (M <= 1024 * 16) f(integral_constant<int, 1024>{}, integral_constant<int, 16>{});
  else fail(MLC_ERR_UNSUPPORTED, std::string(what) + ": row too long");
}

template <class T, class TIN>
void row_r2c(mlc_ctx* ctx, const TIN* in, cpx<T>* out, int Nx, int Ny, const double* shift) {
  const int M = Nx / 2;
  Plan p = plan_for(M, "row FFT");
  const cpx<T>* twM = twiddles<T>(ctx, M);
  const cpx<T>* twN = twiddles<T>(ctx, Nx);
  size_t lds = sizeof(cpx<T>) * M;
  with_row_group(M, "row FFT", [&](auto gc, auto ec) {
    constexpr int G = decltype(gc)::value, EPT = decltype(ec)::value;
    auto k = k_row_r2c<T, TIN, G, EPT>;
    allow_lds(k, lds);
    MLC_LAUNCH(ctx, "fft_row_r2c", k, dim3(Ny), dim3(G), lds, in, out, Nx, shift, p, twM, twN);
  });
}

template <class T, int MODE>
void row_c2r(mlc_ctx* ctx, const cpx<T>* in0, const cpx<T>* in1, T* out, int Nx, int Ny,
             const double* shift) {
  const int M = Nx / 2;
  Plan p = plan_for(M, "row FFT");
  const cpx<T>* twM = twiddles<T>(ctx, M);
  const cpx<T>* twN = twiddles<T>(ctx, Nx);
  size_t lds = sizeof(cpx<T>) * M * (MODE == 2 ? 2 : 1);
  if (lds > kLdsBytes) fail(MLC_ERR_UNSUPPORTED, "row FFT: row does not fit in LDS");
  with_row_group(M, "row FFT", [&](auto gc, auto ec) {
    constexpr int G = decltype(gc)::value, EPT = decltype(ec)::value;
    auto k = k_row_c2r<T, G, EPT, MODE>;
    allow_lds(k, lds);
    MLC_LAUNCH(ctx, MODE == 2 ? "fft_row_c2r_node" : "fft_row_c2r", k, dim3(Ny), dim3(G), lds, in0,
               in1, out, Nx, shift, p, twM, twN);
  });
}

// Column launch geometry: the widest CB (power of two <= 16) whose buffers
// fit in LDS, and the per-thread element bound EPT in {8, 16, 24}.
struct ColGeom {
  int CB, EPT, LS;
  size_t lds;
};

// Column geometries with a kernel instantiation, in order of preference
// (widest CB first, then fewest elements per thread).
constexpr int kColCases[][2] = {{16, 8}, {8, 8}, {8, 16}, {4, 8}, {4, 16}, {2, 8},
                                {2, 16}, {2, 24}, {1, 8}, {1, 16}, {1, 24}};

template <class T>
ColGeom col_geom(int Ny) {
  const int LS = Ny + 1;  // odd stride (in cpx units) between column buffers
  for (auto& cs : kColCases) {
    const int CB = cs[0], EPT = cs[1];
    size_t lds = sizeof(cpx<T>) * (size_t)LS * CB;
    if (lds <= kLdsBytes && Ny * CB <= kColThreads * EPT) return {CB, EPT, LS, lds};
  }
  fail(MLC_ERR_UNSUPPORTED, "column FFT: column of " + std::to_string(Ny) + " does not fit in LDS");
}

template <class F>
void with_col(const ColGeom& g, F&& f) {
#define MLC_COL_CASE(cb, ept)                                                   \
  if (g.CB == cb && g.EPT == ept) {                                             \
    f(std::integral_constant<int, cb>{}, std::integral_constant<int, ept>{});   \
    return;                                                                     \
  }
  MLC_COL_CASE(16, 8) MLC_COL_CASE(8, 8) MLC_COL_CASE(8, 16) MLC_COL_CASE(4, 8)
  MLC_COL_CASE(4, 16) MLC_COL_CASE(2, 8) MLC_COL_CASE(2, 16) MLC_COL_CASE(2, 24)
  MLC_COL_CASE(1, 8) MLC_COL_CASE(1, 16) MLC_COL_CASE(1, 24)
#undef MLC_COL_CASE
  fail(MLC_ERR_UNSUPPORTED, "column FFT: no kernel for this geometry");
}

template <class T, bool INV>
void col_fft(mlc_ctx* ctx, cpx<T>* data, int Ny, int W, T scale) {
  Plan p = plan_for(Ny, "column FFT");
  const cpx<T>* tw = twiddles<T>(ctx, Ny);
  ColGeom g = col_geom<T>(Ny);
  with_col(g, [&](auto cbc, auto eptc) {
    constexpr int CB = decltype(cbc)::value, EPT = decltype(eptc)::value;
    int ngroups = (W + CB - 1) / CB;
    auto k = k_col_fft<T, CB, EPT, INV>;
    allow_lds(k, g.lds);
    MLC_LAUNCH(ctx, INV ? "fft_col_inv" : "fft_col_fwd", k, dim3(((ngroups + 7) / 8) * 8),
               dim3(kColThreads), g.lds, data, Ny, W, scale, p, tw, g.LS);
  });
}

template <class T>
void node_col(mlc_ctx* ctx, const cpx<T>* fmap, const cpx<T>* kr, const cpx<T>* kd, int n,
              cpx<T>* ov, cpx<T>* od, int Ny, int W) {
  Plan p = plan_for(Ny, "column FFT");
  const cpx<T>* tw = twiddles<T>(ctx, Ny);
  ColGeom g = col_geom<T>(Ny);
  with_col(g, [&](auto cbc, auto eptc) {
    constexpr int CB = decltype(cbc)::value, EPT = decltype(eptc)::value;
    int ngroups = (W + CB - 1) / CB;
    auto k = k_node_col<T, CB, EPT>;
    allow_lds(k, g.lds);
    MLC_LAUNCH(ctx, "conv_node_col", k, dim3(((ngroups + 7) / 8) * 8), dim3(kColThreads), g.lds,
               fmap, kr, kd, n, ov, od, Ny, W, p, tw, g.LS);
  });
}

void check_dims(int Nx, int Ny) {
  MLC_REQUIRE(Nx >= 2 && Ny >= 1, "map size must be at least 2 x 1");
  MLC_REQUIRE(Nx % 2 == 0, "map width Nx must be even");
  MLC_REQUIRE((int64_t)Ny * (Nx / 2 + 1) < (int64_t)1 << 30, "map too large for 32-bit offsets");
  (void)plan_for(Nx / 2, "row FFT");
  (void)plan_for(Ny, "column FFT");
}

}  // namespace

// ------------------------------------------------------------------ map handle
struct mlc_map {
  mlc_ctx* ctx = nullptr;
  int precision = MLC_F32;
  int Nx = 0, Ny = 0, W = 0;
  int subtract_mean = 0;
  void* fmap = nullptr;   // Ny x W complex (precision), spectrum / (Nx*Ny)
  double* mean = nullptr; // device FP64: map mean (0 when not subtracted)
  void* g0 = nullptr;     // Ny x W complex scratch (value column output)
  void* g1 = nullptr;     // Ny x W complex scratch (derivative column output)
};

namespace {

size_t elem(int precision) { return precision == MLC_F64 ? sizeof(double) : sizeof(float); }

template <class T>
void map_transform(mlc_map* m, const double* d_map) {
  mlc_ctx* ctx = m->ctx;
  const int64_t npix = (int64_t)m->Nx * m->Ny;
  if (m->subtract_mean) {
    Workspace ws(ctx);
    int np = 1024;
    int ip = ws.add(sizeof(double) * np);
    ws.commit();
    MLC_LAUNCH(ctx, "map_sum_partial", k_sum_partial, dim3(np), dim3(256), 0, d_map, npix,
               ws.get<double>(ip));
    MLC_LAUNCH(ctx, "map_sum_final", k_sum_final, dim3(1), dim3(256), 0, ws.get<double>(ip), np,
               1.0 / (double)npix, m->mean);
  } else {
    hip_check(hipMemsetAsync(m->mean, 0, sizeof(double), ctx->stream), "hipMemsetAsync(mean)");
  }
  cpx<T>* F = (cpx<T>*)m->fmap;
  row_r2c<T, double>(ctx, d_map, F, m->Nx, m->Ny, m->mean);
  col_fft<T, false>(ctx, F, m->Ny, m->W, (T)(1.0 / (double)npix));
}

template <class T>
void convolve_fused(mlc_map* m, const ProfileParams& P, void* d_value, void* d_node) {
  mlc_ctx* ctx = m->ctx;
  const int n = P.n, W = m->W;
  const bool want_d = d_node != nullptr;
  Workspace ws(ctx);
  int ip = ws.add(sizeof(double) * n * n), idp = ws.add(sizeof(double) * n * n);
  int ir = ws.add(sizeof(cpx<T>) * n * W), idr = ws.add(sizeof(cpx<T>) * n * W);
  ws.commit();
  double *prof = ws.get<double>(ip), *dprof = ws.get<double>(idp);
  cpx<T>*R = ws.get<cpx<T>>(ir), *dR = ws.get<cpx<T>>(idr);
  MLC_LAUNCH(ctx, "profile", k_profile, dim3(1), dim3(1024), 0, P, prof, dprof);
  const cpx<double>* tw64 = twiddles<double>(ctx, m->Nx);
  MLC_LAUNCH(ctx, "conv_kernel_rows", k_kernel_rows<T>, dim3(cdiv(W, 256), n), dim3(256), 0, prof,
             dprof, n, m->Nx, W, tw64, R, want_d ? dR : nullptr);
  cpx<T>* G0 = (cpx<T>*)m->g0;
  cpx<T>* G1 = (cpx<T>*)m->g1;
  node_col<T>(ctx, (const cpx<T>*)m->fmap, R, dR, n, G0, want_d ? G1 : nullptr, m->Ny, W);
  if (d_node) row_c2r<T, 2>(ctx, G0, G1, (T*)d_node, m->Nx, m->Ny, m->mean);
  if (d_value) row_c2r<T, 1>(ctx, G0, nullptr, (T*)d_value, m->Nx, m->Ny, m->mean);
}

template <class T>
void convolve_plain(mlc_map* m, const ProfileParams& P, void* d_value, void* d_node) {
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
               (cpx<T>*)d_node, npix);
    sync(ctx);
  }
}

ProfileParams make_params(int shape, double rhalf, double pix, double rin, double a, int Nx, int Ny) {
  MLC_REQUIRE(shape >= 0 && shape <= 5, "unknown profile shape");
  MLC_REQUIRE(std::isfinite(rhalf) && rhalf > 0, "rhalf must be positive and finite");
  MLC_REQUIRE(std::isfinite(pix) && pix > 0, "pixel size must be positive and finite");
  if (shape == MLC_SHAPE_GAUSSIAN_HOLE) MLC_REQUIRE(rin >= 0 && rin < rhalf, "need 0 <= Rin < rhalf");
  if (shape == MLC_SHAPE_WAVY) MLC_REQUIRE(a != 0.0, "wavy profile needs a != 0");
  ProfileParams P;
  MLC_REQUIRE(profile_params(shape, rhalf, pix, rin, a, &P), "profile has an empty grid");
  if (Nx > 0)
    MLC_REQUIRE(P.n <= Nx && P.n <= Ny, "profile grid (" + std::to_string(P.n) +
                                            " px) is larger than the map");
  return P;
}

}  // namespace

namespace mlc {
// used by mlc_expanding_source (mlc_pairs.hip)
void map_convolve_value(mlc_map* m, const ProfileParams& P, void* d_value) {
  if (m->precision == MLC_F64) convolve_fused<double>(m, P, d_value, nullptr);
  else convolve_fused<float>(m, P, d_value, nullptr);
}
ProfileParams checked_params(int shape, double rhalf, double pix, double rin, double a, int Nx,
                             int Ny) {
  return make_params(shape, rhalf, pix, rin, a, Nx, Ny);
}
}  // namespace mlc

// ------------------------------------------------------------------ C ABI
extern "C" {

double mlc_parity_round(double x) {
  std::string s = std::to_string(x);  // "%f"
  return (double)std::stof(s);
}

int mlc_profile_size(int shape, double rhalf, double pix, double rin, double a) {
  ProfileParams P;
  if (!(rhalf > 0) || !(pix > 0) || !profile_params(shape, rhalf, pix, rin, a, &P)) return -1;
  return P.n;
}

int mlc_profile(mlc_ctx* ctx, int shape, double rhalf, double pix, double rin, double a, int* n,
                double* d_prof, double* d_dprof, int64_t cap) {
  return guarded(ctx, [&] {
    ProfileParams P = make_params(shape, rhalf, pix, rin, a, 0, 0);
    MLC_REQUIRE(n, "mlc_profile: null n");
    *n = P.n;
    if (!d_prof && !d_dprof) return;
    MLC_REQUIRE((int64_t)P.n * P.n <= cap, "mlc_profile: capacity too small");
    Workspace ws(ctx);
    int ip = ws.add(sizeof(double) * P.n * P.n), idp = ws.add(sizeof(double) * P.n * P.n);
    ws.commit();
    double* pp = d_prof ? d_prof : ws.get<double>(ip);
    double* dp = d_dprof ? d_dprof : ws.get<double>(idp);
    MLC_LAUNCH(ctx, "profile", k_profile, dim3(1), dim3(1024), 0, P, pp, dp);
  });
}

int mlc_map_create(mlc_ctx* ctx, int precision, const double* d_map, int Nx, int Ny,
                   int subtract_mean, mlc_map** out) {
  mlc_map* m = nullptr;
  int rc = guarded(ctx, [&] {
    MLC_REQUIRE(out, "mlc_map_create: null output");
    *out = nullptr;
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_map_create: bad precision");
    MLC_REQUIRE(d_map, "mlc_map_create: null map");
    check_dims(Nx, Ny);
    m = new mlc_map();
    m->ctx = ctx;
    m->precision = precision;
    m->Nx = Nx;
    m->Ny = Ny;
    m->W = Nx / 2 + 1;
    m->subtract_mean = subtract_mean != 0;
    size_t spec = 2 * elem(precision) * (size_t)Ny * m->W;
    hip_check(hipMalloc(&m->fmap, spec), "hipMalloc(map spectrum)");
    hip_check(hipMalloc(&m->g0, spec), "hipMalloc(column scratch)");
    hip_check(hipMalloc(&m->g1, spec), "hipMalloc(column scratch)");
    hip_check(hipMalloc((void**)&m->mean, sizeof(double)), "hipMalloc(mean)");
    if (precision == MLC_F64) map_transform<double>(m, d_map);
    else map_transform<float>(m, d_map);
    *out = m;
  });
  if (rc != MLC_OK && m) mlc_map_destroy(m);
  return rc;
}

void mlc_map_destroy(mlc_map* m) {
  if (!m) return;
  if (m->ctx) {
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    if (prev >= 0) (void)hipSetDevice(prev);
  }
  (void)hipFree(m->fmap);
  (void)hipFree(m->g0);
  (void)hipFree(m->g1);
  (void)hipFree(m->mean);
  delete m;
}

int mlc_map_update(mlc_map* m, const double* d_map) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(d_map, "mlc_map_update: null map");
    if (m->precision == MLC_F64) map_transform<double>(m, d_map);
    else map_transform<float>(m, d_map);
  });
}

int mlc_map_convolve(mlc_map* m, int shape, double rhalf, double pix, double rin, double a,
                     void* d_value, void* d_node) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(d_value || d_node, "mlc_map_convolve: no output");
    ProfileParams P = make_params(shape, rhalf, pix, rin, a, m->Nx, m->Ny);
    if (m->precision == MLC_F64) convolve_fused<double>(m, P, d_value, d_node);
    else convolve_fused<float>(m, P, d_value, d_node);
  });
}

int mlc_map_convolve_plain(mlc_map* m, int shape, double rhalf, double pix, double rin, double a,
                           void* d_value, void* d_node) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(d_value || d_node, "mlc_map_convolve_plain: no output");
    ProfileParams P = make_params(shape, rhalf, pix, rin, a, m->Nx, m->Ny);
    if (m->precision == MLC_F64) convolve_plain<double>(m, P, d_value, d_node);
    else convolve_plain<float>(m, P, d_value, d_node);
  });
}

int mlc_fft_r2c(mlc_ctx* ctx, int precision, const void* d_in, void* d_out, int Nx, int Ny) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_fft_r2c: bad precision");
    MLC_REQUIRE(d_in && d_out, "mlc_fft_r2c: null pointer");
    check_dims(Nx, Ny);
    const int W = Nx / 2 + 1;
    if (precision == MLC_F64) {
      row_r2c<double, double>(ctx, (const double*)d_in, (cpx<double>*)d_out, Nx, Ny, nullptr);
      col_fft<double, false>(ctx, (cpx<double>*)d_out, Ny, W, 1.0);
    } else {
      row_r2c<float, float>(ctx, (const float*)d_in, (cpx<float>*)d_out, Nx, Ny, nullptr);
      col_fft<float, false>(ctx, (cpx<float>*)d_out, Ny, W, 1.0f);
    }
  });
}

int mlc_fft_c2r(mlc_ctx* ctx, int precision, const void* d_in, void* d_out, int Nx, int Ny) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_fft_c2r: bad precision");
    MLC_REQUIRE(d_in && d_out, "mlc_fft_c2r: null pointer");
    check_dims(Nx, Ny);
    const int W = Nx / 2 + 1;
    size_t bytes = 2 * elem(precision) * (size_t)Ny * W;
    Workspace ws(ctx);
    int it = ws.add(bytes);
    ws.commit();
    void* tmp = ws.get<void>(it);
    hip_check(hipMemcpyAsync(tmp, d_in, bytes, hipMemcpyDeviceToDevice, ctx->stream),
              "hipMemcpyAsync(D2D)");
    if (precision == MLC_F64) {
      col_fft<double, true>(ctx, (cpx<double>*)tmp, Ny, W, 1.0);
      row_c2r<double, 0>(ctx, (const cpx<double>*)tmp, nullptr, (double*)d_out, Nx, Ny, nullptr);
    } else {
      col_fft<float, true>(ctx, (cpx<float>*)tmp, Ny, W, 1.0f);
      row_c2r<float, 0>(ctx, (const cpx<float>*)tmp, nullptr, (float*)d_out, Nx, Ny, nullptr);
    }
  });
}

}  // extern "C"
