// mlc_traj.hip -- per-light-curve magnitude chi2 and its adjoint along
// straight trajectories (the north-star generalisation of kernel_sample,
// gpu_functions.cu:604-620, and of the chi2 of kernel_chi2, 700-750).
//
// A curve c has parameters (x0, y0, theta, v, r0, w, dm), a profile p and an
// observation set o.  For every epoch j of o (band b, time t_j):
//   position  x = x0 + v cos(theta) t_j,  y = y0 + v sin(theta) t_j   (FP64)
//   radius    r = band_scale[b] (r0 + w t_j)
//   mu        bilinear taps (pixel centres at integer coordinates, the
//             convention of rectGrid.cpp:349-378) of the (mu, dmu/drad) node
//             maps of (p, b); cubic Hermite in r between two nodes, a
//             first-order Taylor step outside the node range
//   m         -2.5 log10(mu) + dm                                        (FP64)
//   chi2     += ((mobs_j - m)/sig_j)^2                                   (FP64)
// and the gradient of chi2 with respect to the seven parameters follows by
// the chain rule through the same bilinear weights.
//
// Layout: one wavefront per curve (lanes stride over its epochs, all bands
// concatenated), four curves per workgroup, wavefront reductions at the end.
// A curve that leaves the pixel-centre hull or meets mu <= 0 gets chi2 = +inf
// and a zero gradient.
#include <cmath>
#include <vector>

#include "mlc_fft_core.hpp"
#include "mlc_internal.hpp"

using namespace mlc;

namespace {

constexpr double kLn10 = 2.30258509299404568401799145468436421;

struct NodeTab {
  const int* off;     // P*B
  const int* cnt;     // P*B
  const double* rad;  // nnodes
  const double* bs;   // B
};

template <class T>
struct Tap {
  T v, dx, dy;  // bilinear value and its x / y derivative
};

template <class T>
__device__ __forceinline__ void taps(const cpx<T>* __restrict__ m, int Nx, T fx, T fy, Tap<T>* M,
                                     Tap<T>* D) {
  const cpx<T> a = m[0], b = m[1], c = m[Nx], d = m[Nx + 1];
  const T gx = (T)1 - fx, gy = (T)1 - fy;
  M->v = gy * (gx * a.x + fx * b.x) + fy * (gx * c.x + fx * d.x);
  M->dx = gy * (b.x - a.x) + fy * (d.x - c.x);
  M->dy = gx * (c.x - a.x) + fx * (d.x - b.x);
  D->v = gy * (gx * a.y + fx * b.y) + fy * (gx * c.y + fx * d.y);
  D->dx = gy * (b.y - a.y) + fy * (d.y - c.y);
  D->dy = gx * (c.y - a.y) + fx * (d.y - b.y);
}

// mu and its derivatives with respect to x, y and r at one epoch.
struct Sample {
  double mu, mx, my, mr;
  bool ok;
};

template <class T>
__device__ __forceinline__ Sample sample_at(const cpx<T>* __restrict__ nodes, size_t mapsz, int Nx,
                                            int Ny, const NodeTab& tab, int pb, double x, double y,
                                            double r) {
  Sample s;
  s.ok = (x >= 0.0 && x <= (double)(Nx - 1) && y >= 0.0 && y <= (double)(Ny - 1));
  if (!s.ok) {
    s.mu = s.mx = s.my = s.mr = 0.0;
    return s;
  }
  int ix = min((int)x, Nx - 2), iy = min((int)y, Ny - 2);
  const T fx = (T)(x - ix), fy = (T)(y - iy);
  const size_t pix = (size_t)iy * Nx + ix;
  const int noff = tab.off[pb], ncnt = tab.cnt[pb];
  int k = 0;
  bool taylor = true;
  double sr = 0.0, h = 1.0;
  if (ncnt >= 2) {
    const double rl = tab.rad[noff], rh = tab.rad[noff + ncnt - 1];
    if (r <= rl) k = 0;
    else if (r >= rh) k = ncnt - 1;
    else {
      taylor = false;
      while (k < ncnt - 2 && tab.rad[noff + k + 1] <= r) ++k;
      h = tab.rad[noff + k + 1] - tab.rad[noff + k];
      sr = (r - tab.rad[noff + k]) / h;
    }
  }
  Tap<T> M0, D0;
  taps<T>(nodes + mapsz * (noff + k) + pix, Nx, fx, fy, &M0, &D0);
  if (taylor) {
    const T dr = (T)(r - tab.rad[noff + k]);
    s.mu = (double)(M0.v + D0.v * dr);
    s.mx = (double)(M0.dx + D0.dx * dr);
    s.my = (double)(M0.dy + D0.dy * dr);
    s.mr = (double)D0.v;
  } else {
    Tap<T> M1, D1;
    taps<T>(nodes + mapsz * (noff + k + 1) + pix, Nx, fx, fy, &M1, &D1);
    const T S = (T)sr, S2 = S * S, S3 = S2 * S, H = (T)h;
    const T h00 = 2 * S3 - 3 * S2 + 1, h10 = (S3 - 2 * S2 + S) * H;
    const T h01 = -2 * S3 + 3 * S2, h11 = (S3 - S2) * H;
    const T d00 = 6 * S2 - 6 * S, d10 = (3 * S2 - 4 * S + 1) * H;
    const T d01 = -6 * S2 + 6 * S, d11 = (3 * S2 - 2 * S) * H;
    s.mu = (double)(h00 * M0.v + h10 * D0.v + h01 * M1.v + h11 * D1.v);
    s.mx = (double)(h00 * M0.dx + h10 * D0.dx + h01 * M1.dx + h11 * D1.dx);
    s.my = (double)(h00 * M0.dy + h10 * D0.dy + h01 * M1.dy + h11 * D1.dy);
    s.mr = (double)(d00 * M0.v + d10 * D0.v + d01 * M1.v + d11 * D1.v) / h;
  }
  s.ok = s.mu > 0.0;
  return s;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <class T>
__global__ __launch_bounds__(256) void k_lc_chi2_grad(
    const cpx<T>* __restrict__ nodes, int Nx, int Ny, int P, int B, NodeTab tab, int ncurves,
    const double* __restrict__ params, const int* __restrict__ prof, const int* __restrict__ obs,
    int nobs, const int* __restrict__ obs_off, int nepoch, const double* __restrict__ tt,
    const double* __restrict__ mobs, const double* __restrict__ sig, double* __restrict__ chi2,
    double* __restrict__ grad) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncurves) return;  // whole wavefront
  const size_t mapsz = (size_t)Nx * Ny;
  const double* Pc = params + (size_t)c * 7;
  const double x0 = Pc[0], y0 = Pc[1], th = Pc[2], v = Pc[3], r0 = Pc[4], w = Pc[5], dm = Pc[6];
  const double ct = cos(th), st = sin(th);
  const int p = prof[c], o = obs[c];
  double acc = 0.0, g[7] = {0, 0, 0, 0, 0, 0, 0};
  bool bad = !(p >= 0 && p < P && o >= 0 && o < nobs);
  if (!bad) {
    const int* off = obs_off + (size_t)o * B;
    const int e0 = off[0], e1 = off[B];
    bad = !(e0 >= 0 && e1 <= nepoch && e0 <= e1);
    if (!bad) {
      for (int j = e0 + lane; j < e1; j += 64) {
        int b = 0;
        while (b < B - 1 && j >= off[b + 1]) ++b;
        const double te = tt[j];
        const double x = x0 + v * ct * te, y = y0 + v * st * te;
        const double bsc = tab.bs[b];
        const double r = bsc * (r0 + w * te);
        Sample s = sample_at<T>(nodes, mapsz, Nx, Ny, tab, p * B + b, x, y, r);
        if (!s.ok) {
          bad = true;
          break;
        }
        const double sg = sig[j];
        const double m = -2.5 * log10(s.mu) + dm;
        const double rho = (mobs[j] - m) / sg;
        acc += rho * rho;
        const double dc = 2.0 * rho * 2.5 / (kLn10 * s.mu * sg);
        g[0] += dc * s.mx;
        g[1] += dc * s.my;
        g[2] += dc * (s.mx * (-v * st * te) + s.my * (v * ct * te));
        g[3] += dc * (s.mx * ct * te + s.my * st * te);
        g[4] += dc * s.mr * bsc;
        g[5] += dc * s.mr * bsc * te;
        g[6] += -2.0 * rho / sg;
      }
    }
  }
  const bool any_bad = __ballot(bad) != 0;
  acc = wave_sum(acc);
#pragma unroll
  for (int q = 0; q < 7; ++q) g[q] = wave_sum(g[q]);
  if (lane == 0) {
    chi2[c] = any_bad ? INFINITY : acc;
#pragma unroll
    for (int q = 0; q < 7; ++q) grad[(size_t)c * 7 + q] = any_bad ? 0.0 : g[q];
  }
}

// ---- plain path: stage 1 writes (mu, dmu/dx, dmu/dy, dmu/dr) per sample,
// stage 2 reduces them per curve in epoch order.
template <class T>
__global__ __launch_bounds__(256) void k_lc_samples(
    const cpx<T>* __restrict__ nodes, int Nx, int Ny, int P, int B, NodeTab tab, int ncurves,
    const double* __restrict__ params, const int* __restrict__ prof, const int* __restrict__ obs,
    int nobs, const int* __restrict__ obs_off, int nepoch, const double* __restrict__ tt, int smax,
    double* __restrict__ samp) {
  const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= (int64_t)ncurves * smax) return;
  const int c = (int)(id / smax), k = (int)(id % smax);
  const int p = prof[c], o = obs[c];
  double* out = samp + (size_t)id * 4;
  if (!(p >= 0 && p < P && o >= 0 && o < nobs)) {
    out[0] = NAN;
    return;
  }
  const int* off = obs_off + (size_t)o * B;
  const int j = off[0] + k;
  if (j >= off[B]) return;
  if (!(off[0] >= 0 && off[B] <= nepoch)) {
    out[0] = NAN;
    return;
  }
  int b = 0;
  while (b < B - 1 && j >= off[b + 1]) ++b;
  const double* Pc = params + (size_t)c * 7;
  const double te = tt[j];
  const double x = Pc[0] + Pc[3] * cos(Pc[2]) * te, y = Pc[1] + Pc[3] * sin(Pc[2]) * te;
  const double r = tab.bs[b] * (Pc[4] + Pc[5] * te);
  Sample s = sample_at<T>(nodes, (size_t)Nx * Ny, Nx, Ny, tab, p * B + b, x, y, r);
  out[0] = s.ok ? s.mu : NAN;
  out[1] = s.mx;
  out[2] = s.my;
  out[3] = s.mr;
}

__global__ __launch_bounds__(256) void k_lc_reduce(int B, const double* __restrict__ bs, int ncurves,
                                                   const double* __restrict__ params,
                                                   const int* __restrict__ obs, int nobs,
                                                   const int* __restrict__ obs_off, int nepoch,
                                                   const double* __restrict__ tt,
                                                   const double* __restrict__ mobs,
                                                   const double* __restrict__ sig, int smax,
                                                   const double* __restrict__ samp,
                                                   double* __restrict__ chi2, double* __restrict__ grad) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncurves) return;
  const double* Pc = params + (size_t)c * 7;
  const double v = Pc[3], ct = cos(Pc[2]), st = sin(Pc[2]), dm = Pc[6];
  double acc = 0.0, g[7] = {0, 0, 0, 0, 0, 0, 0};
  const double* sp = samp + (size_t)c * smax * 4;
  const int o = obs[c];
  bool bad = !(o >= 0 && o < nobs);
  const int* off = obs_off + (size_t)(bad ? 0 : o) * B;
  bad = bad || !(off[0] >= 0 && off[B] <= nepoch && off[0] <= off[B]);
  const int n = bad ? 0 : off[B] - off[0];
  if (n > 0 && !(sp[0] == sp[0])) bad = true;
  for (int k = 0; k < n && !bad; ++k) {
    const int j = off[0] + k;
    int b = 0;
    while (b < B - 1 && j >= off[b + 1]) ++b;
    const double mu = sp[4 * k], mx = sp[4 * k + 1], my = sp[4 * k + 2], mr = sp[4 * k + 3];
    if (!(mu > 0.0)) {
      bad = true;
      break;
    }
    const double te = tt[j], sg = sig[j];
    const double m = -2.5 * log10(mu) + dm;
    const double rho = (mobs[j] - m) / sg;
    acc += rho * rho;
    const double dc = 2.0 * rho * 2.5 / (kLn10 * mu * sg);
    g[0] += dc * mx;
    g[1] += dc * my;
    g[2] += dc * (mx * (-v * st * te) + my * (v * ct * te));
    g[3] += dc * (mx * ct * te + my * st * te);
    g[4] += dc * mr * bs[b];
    g[5] += dc * mr * bs[b] * te;
    g[6] += -2.0 * rho / sg;
  }
  chi2[c] = bad ? INFINITY : acc;
  for (int q = 0; q < 7; ++q) grad[(size_t)c * 7 + q] = bad ? 0.0 : g[q];
}

__global__ void k_max_span(const int* __restrict__ obs_off, int nobs, int B, int* __restrict__ out) {
  int m = 0;
  for (int o = threadIdx.x; o < nobs; o += blockDim.x) m = max(m, obs_off[(size_t)(o + 1) * B] - obs_off[(size_t)o * B]);
  atomicMax(out, m);
}

struct Args {
  int precision;
  const void* nodes;
  int nnodes, Nx, Ny, P, B;
  const int *node_off, *node_cnt;
  const double *node_rad, *band_scale;
  int ncurves;
  const double* params;
  const int *prof, *obs;
  int nobs;
  const int* obs_off;
  int nepoch;
  const double *t, *mobs, *sig;
  double *chi2, *grad;
};

void validate(const Args& a) {
  MLC_REQUIRE(a.precision == MLC_F32 || a.precision == MLC_F64, "lc_chi2_grad: bad precision");
  MLC_REQUIRE(a.Nx >= 2 && a.Ny >= 2 && a.P >= 1 && a.B >= 1 && a.nnodes >= 1 && a.ncurves >= 0 &&
                  a.nobs >= 1 && a.nepoch >= 0,
              "lc_chi2_grad: bad sizes");
  MLC_REQUIRE(a.nodes && a.node_off && a.node_cnt && a.node_rad && a.band_scale,
              "lc_chi2_grad: null node table");
  MLC_REQUIRE(a.ncurves == 0 || (a.params && a.prof && a.obs && a.obs_off && a.chi2 && a.grad),
              "lc_chi2_grad: null curve arrays");
  MLC_REQUIRE(a.nepoch == 0 || (a.t && a.mobs && a.sig), "lc_chi2_grad: null epoch arrays");
  for (int i = 0; i < a.P * a.B; ++i) {
    MLC_REQUIRE(a.node_cnt[i] >= 1 && a.node_off[i] >= 0 && a.node_off[i] + a.node_cnt[i] <= a.nnodes,
                "lc_chi2_grad: node table entry out of range");
    for (int k = 1; k < a.node_cnt[i]; ++k)
      MLC_REQUIRE(a.node_rad[a.node_off[i] + k] > a.node_rad[a.node_off[i] + k - 1],
                  "lc_chi2_grad: node radii must increase");
  }
}

NodeTab upload_tab(mlc_ctx* ctx, const Args& a, Workspace& ws, int io, int ic, int ir, int ib) {
  h2d(ctx, ws.get<void>(io), a.node_off, sizeof(int) * a.P * a.B);
  h2d(ctx, ws.get<void>(ic), a.node_cnt, sizeof(int) * a.P * a.B);
  h2d(ctx, ws.get<void>(ir), a.node_rad, sizeof(double) * a.nnodes);
  h2d(ctx, ws.get<void>(ib), a.band_scale, sizeof(double) * a.B);
  return NodeTab{ws.get<int>(io), ws.get<int>(ic), ws.get<double>(ir), ws.get<double>(ib)};
}

void run_fused(mlc_ctx* ctx, const Args& a) {
  validate(a);
  if (a.ncurves == 0) return;
  Workspace ws(ctx);
  int io = ws.add(sizeof(int) * a.P * a.B), ic = ws.add(sizeof(int) * a.P * a.B),
      ir = ws.add(sizeof(double) * a.nnodes), ib = ws.add(sizeof(double) * a.B);
  ws.commit();
  NodeTab tab = upload_tab(ctx, a, ws, io, ic, ir, ib);
  dim3 grid(cdiv(a.ncurves, 4));
  if (a.precision == MLC_F64)
    MLC_LAUNCH(ctx, "lc_chi2_grad", k_lc_chi2_grad<double>, grid, dim3(256), 0,
               (const cpx<double>*)a.nodes, a.Nx, a.Ny, a.P, a.B, tab, a.ncurves, a.params, a.prof,
               a.obs, a.nobs, a.obs_off, a.nepoch, a.t, a.mobs, a.sig, a.chi2, a.grad);
  else
    MLC_LAUNCH(ctx, "lc_chi2_grad", k_lc_chi2_grad<float>, grid, dim3(256), 0,
               (const cpx<float>*)a.nodes, a.Nx, a.Ny, a.P, a.B, tab, a.ncurves, a.params, a.prof,
               a.obs, a.nobs, a.obs_off, a.nepoch, a.t, a.mobs, a.sig, a.chi2, a.grad);
  sync(ctx);  // the node tables were staged from host memory of this call
}

void run_plain(mlc_ctx* ctx, const Args& a) {
  validate(a);
  if (a.ncurves == 0) return;
  Workspace ws(ctx);
  int io = ws.add(sizeof(int) * a.P * a.B), ic = ws.add(sizeof(int) * a.P * a.B),
      ir = ws.add(sizeof(double) * a.nnodes), ib = ws.add(sizeof(double) * a.B),
      im = ws.add(sizeof(int));
  ws.commit();
  NodeTab tab = upload_tab(ctx, a, ws, io, ic, ir, ib);
  hip_check(hipMemsetAsync(ws.get<void>(im), 0, sizeof(int), ctx->stream), "hipMemsetAsync");
  MLC_LAUNCH(ctx, "lc_plain_span", k_max_span, dim3(1), dim3(256), 0, a.obs_off, a.nobs, a.B,
             ws.get<int>(im));
  int smax = 0;
  d2h(ctx, &smax, ws.get<void>(im), sizeof(int));
  sync(ctx);
  smax = std::max(smax, 1);
  const size_t sbytes = sizeof(double) * 4 * (size_t)a.ncurves * smax;
  void* samp = nullptr;
  hip_check(hipMalloc(&samp, sbytes), "hipMalloc(samples)");
  try {
    const int64_t ns = (int64_t)a.ncurves * smax;
    if (a.precision == MLC_F64)
      MLC_LAUNCH(ctx, "lc_plain_samples", k_lc_samples<double>, dim3(cdiv(ns, 256)), dim3(256), 0,
                 (const cpx<double>*)a.nodes, a.Nx, a.Ny, a.P, a.B, tab, a.ncurves, a.params, a.prof,
                 a.obs, a.nobs, a.obs_off, a.nepoch, a.t, smax, (double*)samp);
    else
      MLC_LAUNCH(ctx, "lc_plain_samples", k_lc_samples<float>, dim3(cdiv(ns, 256)), dim3(256), 0,
                 (const cpx<float>*)a.nodes, a.Nx, a.Ny, a.P, a.B, tab, a.ncurves, a.params, a.prof,
                 a.obs, a.nobs, a.obs_off, a.nepoch, a.t, smax, (double*)samp);
    sync(ctx);
    MLC_LAUNCH(ctx, "lc_plain_reduce", k_lc_reduce, dim3(cdiv(a.ncurves, 256)), dim3(256), 0, a.B,
               tab.bs, a.ncurves, a.params, a.obs, a.nobs, a.obs_off, a.nepoch, a.t, a.mobs, a.sig,
               smax,
               (const double*)samp, a.chi2, a.grad);
    sync(ctx);
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(samp);
    throw;
  }
  hip_check(hipFree(samp), "hipFree(samples)");
}

}  // namespace

extern "C" {

int mlc_lc_chi2_grad(mlc_ctx* ctx, int precision, const void* d_nodes, int nnodes, int Nx, int Ny,
                     int P, int B, const int* node_off, const int* node_cnt, const double* node_rad,
                     const double* band_scale, int ncurves, const double* d_params, const int* d_prof,
                     const int* d_obs, int nobs, const int* d_obs_off, int nepoch, const double* d_t,
                     const double* d_mobs, const double* d_sig, double* d_chi2, double* d_grad) {
  return guarded(ctx, [&] {
    Args a{precision, d_nodes, nnodes, Nx, Ny, P, B, node_off, node_cnt, node_rad, band_scale,
           ncurves, d_params, d_prof, d_obs, nobs, d_obs_off, nepoch, d_t, d_mobs, d_sig, d_chi2,
           d_grad};
    run_fused(ctx, a);
  });
}

int mlc_lc_chi2_grad_plain(mlc_ctx* ctx, int precision, const void* d_nodes, int nnodes, int Nx,
                           int Ny, int P, int B, const int* node_off, const int* node_cnt,
                           const double* node_rad, const double* band_scale, int ncurves,
                           const double* d_params, const int* d_prof, const int* d_obs, int nobs,
                           const int* d_obs_off, int nepoch, const double* d_t, const double* d_mobs,
                           const double* d_sig, double* d_chi2, double* d_grad) {
  return guarded(ctx, [&] {
    Args a{precision, d_nodes, nnodes, Nx, Ny, P, B, node_off, node_cnt, node_rad, band_scale,
           ncurves, d_params, d_prof, d_obs, nobs, d_obs_off, nepoch, d_t, d_mobs, d_sig, d_chi2,
           d_grad};
    run_plain(ctx, a);
  });
}

}  // extern "C"
