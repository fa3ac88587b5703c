// mlc_traj.hip -- per-light-curve magnitude chi2 and its adjoint along
// straight trajectories (a generalisation of kernel_sample,
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
// Layout.  The node maps are in 4 x 4 pixel tiles (mlc_tiles.hpp).  The
// fused path first partitions the curves by profile into 64-byte records
// (origin, velocity components, radius law, offset), then runs one launch
// per (profile p, band b): a group of lanes per curve strides over the
// band-b epochs, and the group's FP64 sums go to the curve's accumulator
// (the first band writes it, later bands add in band order).  Only the node
// maps of (p, b) are read during a launch, so they stay in the last-level
// cache; every accumulator has one writer per launch, so the result does not
// depend on scheduling.  A curve that leaves the pixel-centre hull or meets
// mu <= 0 gets chi2 = +inf and a zero gradient.
#include <cmath>
#include <functional>
#include <mutex>
#include <vector>

#include "mlc_fft.hpp"
#include "mlc_fft_core.hpp"
#include "mlc_internal.hpp"
#include "mlc_math.hpp"
#include "mlc_tiles.hpp"

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

// Bilinear taps at pixel (ix, iy)..(ix+1, iy+1) of a tiled node map.
template <class T>
__device__ __forceinline__ void taps(const cpx<T>* __restrict__ m, int Nx, int ix, int iy, T fx,
                                     T fy, Tap<T>* M, Tap<T>* D) {
  const long long o00 = tile_offset(ix, iy, Nx);
  const long long o01 = (ix & 3) != 3 ? o00 + 1 : tile_offset(ix + 1, iy, Nx);
  const long long o10 = (iy & 3) != 3 ? o00 + 4 : tile_offset(ix, iy + 1, Nx);
  const long long o11 = (ix & 3) != 3 ? o10 + 1 : tile_offset(ix + 1, iy + 1, Nx);
  const cpx<T> a = m[o00], b = m[o01], c = m[o10], d = m[o11];
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
  taps<T>(nodes + mapsz * (noff + k), Nx, ix, iy, fx, fy, &M0, &D0);
  if (taylor) {
    const T dr = (T)(r - tab.rad[noff + k]);
    s.mu = (double)(M0.v + D0.v * dr);
    s.mx = (double)(M0.dx + D0.dx * dr);
    s.my = (double)(M0.dy + D0.dy * dr);
    s.mr = (double)D0.v;
  } else {
    Tap<T> M1, D1;
    taps<T>(nodes + mapsz * (noff + k + 1), Nx, ix, iy, fx, fy, &M1, &D1);
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

// ---- fused path
// Curves are grouped by profile (counting partition, one bucket per profile
// plus bucket P for invalid profile indices) into 64-byte records in bucket
// order; the band launches and the finish kernel address the per-curve
// accumulators by record position.
struct alignas(64) CurveRec {
  double x0, y0, vct, vst, r0, w, dm;  // vct, vst = v cos(theta), v sin(theta)
  int c, o;                            // curve index, observation set
};
static_assert(sizeof(CurveRec) == 64, "one record per 64 bytes");

constexpr int kPartChunk = 2048;    // curves per partition block
constexpr int kPartThreads = 1024;  // threads of a partition block (2 curves each)
constexpr int kMaxBuckets = 8192;   // partition bucket counters live in LDS
constexpr int kMaxBandNodes = 16;   // radius nodes of one (profile, band), kernel argument
constexpr int kDefaultLanes = 16;   // lanes per curve when no tuned shape applies

// Buckets of the partition: profile p and a coarse map cell of the curve's
// position at the middle of its epochs, cells in Morton order, so that
// curves processed together read nearby parts of the node maps; bucket P*C
// collects curves with an invalid profile index.
// With several maps (map != NULL: per-curve map index, msz: per-map Nx, Ny)
// the bucket's first key is map * P + profile, and the cell grid spans each
// curve's own map.
struct PartGeom {
  int P, C, L;       // profiles, cells (4^L), bits per cell axis
  double sx, sy;     // cells per pixel along x and y (one map)
  const int* prof;
  const double* params;
  const int* obs;
  int nobs, B, nepoch;
  const int* obs_off;
  const double* tt;
  int nmaps = 1;               // maps; buckets of (map, profile): NP = nmaps * P
  const int* map = nullptr;    // per-curve map index, or NULL (one map)
  const int* msz = nullptr;    // per-map (Nx, Ny)
  __host__ __device__ int NP() const { return nmaps * P; }
};

__device__ __forceinline__ int curve_bucket(const PartGeom& g, int c) {
  const int p = g.prof[c];
  const int mi = g.map ? g.map[c] : 0;
  if (!(p >= 0 && p < g.P && mi >= 0 && mi < g.nmaps)) return g.NP() * g.C;
  double tm = 0.0;
  const int o = g.obs[c];
  if (o >= 0 && o < g.nobs) {
    const int e0 = g.obs_off[(size_t)o * g.B], e1 = g.obs_off[(size_t)o * g.B + g.B];
    if (e0 >= 0 && e1 <= g.nepoch && e0 < e1) tm = 0.5 * (g.tt[e0] + g.tt[e1 - 1]);
  }
  const double* Pc = g.params + (size_t)c * 7;
  double sn, cs;
  sincos(Pc[2], &sn, &cs);
  const int side = 1 << g.L;
  const double sx = g.map ? (double)side / g.msz[2 * mi] : g.sx;
  const double sy = g.map ? (double)side / g.msz[2 * mi + 1] : g.sy;
  const double x = (Pc[0] + Pc[3] * cs * tm) * sx, y = (Pc[1] + Pc[3] * sn * tm) * sy;
  const int cx = (x >= 0.0) ? (int)min(x, (double)(side - 1)) : 0;  // NaN -> 0
  const int cy = (y >= 0.0) ? (int)min(y, (double)(side - 1)) : 0;
  int m = 0;
  for (int bit = 0; bit < g.L; ++bit) m |= (((cx >> bit) & 1) << (2 * bit)) | (((cy >> bit) & 1) << (2 * bit + 1));
  return (mi * g.P + p) * g.C + m;
}

// Per-block bucket counts, stored block-major: bh[block * nq + q].
__global__ __launch_bounds__(kPartThreads) void k_part_count(PartGeom g, int n, int* __restrict__ bh) {
  extern __shared__ int h[];
  const int nq = g.NP() * g.C + 1;
  for (int q = threadIdx.x; q < nq; q += kPartThreads) h[q] = 0;
  __syncthreads();
  const int c0 = blockIdx.x * kPartChunk, c1 = min(n, c0 + kPartChunk);
  for (int c = c0 + threadIdx.x; c < c1; c += kPartThreads) atomicAdd(&h[curve_bucket(g, c)], 1);
  __syncthreads();
  for (int q = threadIdx.x; q < nq; q += kPartThreads) bh[(size_t)blockIdx.x * nq + q] = h[q];
}

// Exclusive prefix of every bucket's counts over the partition blocks, in
// place, and the bucket totals.  A workgroup covers 64 buckets; its 16 lane
// groups take 16 contiguous runs of blocks, whose sums are scanned in LDS.
constexpr int kColScanQ = 64, kColScanSeg = 16;
__global__ __launch_bounds__(kColScanQ * kColScanSeg) void k_part_colscan(int* __restrict__ bh, int nq,
                                                                           int nblk,
                                                                           int* __restrict__ tot) {
  __shared__ int s[kColScanSeg][kColScanQ];
  const int qq = threadIdx.x % kColScanQ, sg = threadIdx.x / kColScanQ;
  const int q = blockIdx.x * kColScanQ + qq;
  const int len = (nblk + kColScanSeg - 1) / kColScanSeg;
  const int b0 = min(nblk, sg * len), b1 = min(nblk, b0 + len);
  int sum = 0;
  if (q < nq)
    for (int b = b0; b < b1; ++b) sum += bh[(size_t)b * nq + q];
  s[sg][qq] = sum;
  __syncthreads();
  int run = 0;
  for (int k = 0; k < sg; ++k) run += s[k][qq];
  if (q < nq) {
    for (int b = b0; b < b1; ++b) {
      const int v = bh[(size_t)b * nq + q];
      bh[(size_t)b * nq + q] = run;
      run += v;
    }
    if (sg == kColScanSeg - 1) tot[q] = run;
  }
}

// Exclusive scan of the nq bucket totals (one workgroup, nq <= 8192) into
// base[], then start[p] / cnt[p] of every profile (p = P: the invalid bucket).
__global__ __launch_bounds__(1024) void k_part_base(const int* __restrict__ tot, int nq, int P, int C,
                                                    int n, int* __restrict__ base,
                                                    int* __restrict__ start, int* __restrict__ cnt) {
  __shared__ int s[1024];
  const int t = threadIdx.x, seg = (nq + 1023) / 1024;
  const int a = min(nq, t * seg), e = min(nq, a + seg);
  int sum = 0;
  for (int i = a; i < e; ++i) sum += tot[i];
  s[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = s[t] - sum;
  for (int i = a; i < e; ++i) {
    base[i] = run;
    run += tot[i];
  }
  __syncthreads();
  for (int p = t; p <= P; p += 1024) {
    const int st = base[p * C];
    start[p] = st;
    cnt[p] = (p < P ? base[(p + 1) * C] : n) - st;
  }
}

// Records of the block's curves at the block's bucket offsets (the order
// inside one block's share of a bucket is whatever the LDS counter gives; no
// result depends on it).
__global__ __launch_bounds__(kPartThreads) void k_part_fill(PartGeom g, int n, const int* __restrict__ bh,
                                                   const int* __restrict__ base,
                                                   CurveRec* __restrict__ rec,
                                                   int* __restrict__ pos_of) {
  extern __shared__ int h[];
  const int nq = g.NP() * g.C + 1;
  for (int q = threadIdx.x; q < nq; q += kPartThreads) h[q] = base[q] + bh[(size_t)blockIdx.x * nq + q];
  __syncthreads();
  const int c0 = blockIdx.x * kPartChunk, c1 = min(n, c0 + kPartChunk);
  for (int c = c0 + threadIdx.x; c < c1; c += kPartThreads) {
    const int pos = atomicAdd(&h[curve_bucket(g, c)], 1);
    const double* Pc = g.params + (size_t)c * 7;
    double sn, cs;
    sincos(Pc[2], &sn, &cs);
    CurveRec r;
    r.x0 = Pc[0];
    r.y0 = Pc[1];
    r.vct = Pc[3] * cs;
    r.vst = Pc[3] * sn;
    r.r0 = Pc[4];
    r.w = Pc[5];
    r.dm = Pc[6];
    r.c = c;
    r.o = g.obs[c];
    rec[pos] = r;
    pos_of[c] = pos;
  }
}

// One epoch as the band kernels read it: time, observed magnitude, 1/sigma.
struct alignas(32) Epoch {
  double t, m, is, pad;
};
constexpr int kStageEpochs = 128;  // epochs per band a wavefront stages in LDS

// The epoch records, and per (observation set o, band b) segment the
// earliest and latest epoch time (NaN when the segment holds a NaN time).
// One wavefront per segment; a segment whose range is invalid is skipped,
// the band kernels flag its curves before reading anything of it.
__global__ __launch_bounds__(256) void k_epoch_prep(const int* __restrict__ obs_off, int nseg,
                                                    int nepoch, const double* __restrict__ tt,
                                                    const double* __restrict__ mobs,
                                                    const double* __restrict__ sig,
                                                    Epoch* __restrict__ ep, double2* __restrict__ rng) {
  const int sgi = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (sgi >= nseg) return;  // uniform per wave
  const int e0 = obs_off[sgi], e1 = obs_off[sgi + 1];
  if (!(e0 >= 0 && e1 <= nepoch && e0 <= e1)) return;
  double lo = INFINITY, hi = -INFINITY;
  bool nan = false;
  for (int j = e0 + lane; j < e1; j += 64) {
    const double t = tt[j];
    ep[j] = Epoch{t, mobs[j], 1.0 / sig[j], 0.0};
    nan |= t != t;
    lo = fmin(lo, t);
    hi = fmax(hi, t);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  const bool any_nan = __ballot(nan) != 0;
  if (lane == 0) rng[sgi] = any_nan ? make_double2(NAN, NAN) : make_double2(lo, hi);
}

// The radius nodes of one (profile, band), passed by value.
struct BandNodes {
  int cnt;                     // >= 1
  double bsc;                  // band_scale[b]
  double rad[kMaxBandNodes];   // ascending
};

// 32-bit pair offset of pixel (x, y) in a tiled map (Nx * Ny < 2^31).
__device__ __forceinline__ int tile32(int x, int y, int Nx) {
  return (((y >> 2) * (Nx >> 2) + (x >> 2)) << 4) + ((y & 3) << 2) + (x & 3);
}

// Sum of v over the LPC lanes of a group, 8 values at once: three halving
// exchanges leave lane li with value index (li / (LPC/8)) & 7, the remaining
// exchanges finish that value.  Returns the lane's value.
template <int LPC>
__device__ __forceinline__ double group_sum8(double (&v)[8], int li) {
  constexpr int o1 = LPC / 2, o2 = LPC / 4, o3 = LPC / 8;
  double a[4], c[2];
  {
    const bool up = li & o1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double send = up ? v[q] : v[q + 4], keep = up ? v[q + 4] : v[q];
      a[q] = keep + __shfl_xor(send, o1, 64);
    }
  }
  {
    const bool up = li & o2;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const double send = up ? a[q] : a[q + 2], keep = up ? a[q + 2] : a[q];
      c[q] = keep + __shfl_xor(send, o2, 64);
    }
  }
  const bool up = li & o3;
  double r = (up ? c[1] : c[0]) + __shfl_xor(up ? c[0] : c[1], o3, 64);
#pragma unroll
  for (int o = o3 / 2; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
  return r;
}

// Radius nodes of one (profile, band) as a sampler reads them: nn >= 1
// ascending radii rad[], 1/(rad[k+1] - rad[k]) in ih[] (FP32 maps), node map
// k at node(k).
template <class NodeAt>
struct NodeView {
  int nn;
  double rlo, rhi;
  const double* rad;
  const double* ih;
  NodeAt node;
};

// One epoch of a curve between its tap loads and its sums: the epoch
// (time te, observed magnitude em, 1/sigma eis), the bilinear weights and
// the raw corner data of the node maps as loaded.  Splitting a sample there
// lets a lane issue the loads of its next epoch before it finishes this one.
typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
template <class T>
struct SampleTaps {
  double te, em, eis;
  T fx, fy;
  T dr;           // single node: r - node radius
  T w[8];         // several nodes: mu and dmu/dr weights of (Ma, Da, Mb, Db), see lc_fetch
  f4a8 top, bot;  // FP32 single node: 16-byte loads of the cell's rows
  bool cross;     // FP32 single node: the cell crosses a tile column
  cpx<T> c[8];    // FP32 single node: c[1], c[3] the right taps of a crossing cell;
                  // FP64 single node: the four corners; several nodes: Ma[0..3], Mb[0..3]
};

// The first half of a sample of curve R at epoch (te, em, eis) on the node
// maps of one (profile, band): position, radius, bilinear weights, the
// node interval and its Hermite weights, and the tap loads.  The caller has
// checked that the position lies in the pixel-centre hull (0 <= x <= Nx - 1,
// 0 <= y <= Ny - 1).  ONE: a single radius node, first-order Taylor step in
// r, taps through 32-bit offsets from node(0) (the map is below 4 GiB).
// rowstep = 4 Nx - 12, the pair-offset step from pixel row 4k+3 to 4k+4.
// lc_fetch and lc_accum fuse only where they call fma: their rounding does
// not depend on how the compiler schedules the code around them.
template <class T, bool ONE, class NodeAt>
__device__ __forceinline__ SampleTaps<T> lc_fetch(const NodeView<NodeAt>& nv, int Nx, int Ny, int rowstep,
                                                  double bsc, const CurveRec& R, double te, double em,
                                                  double eis) {
#pragma clang fp contract(off)
  constexpr bool kF32 = sizeof(T) == 4;
  SampleTaps<T> P;
  P.te = te;
  P.em = em;
  P.eis = eis;
  const double x = fma(R.vct, te, R.x0), y = fma(R.vst, te, R.y0);
  const double rt = fma(R.w, te, R.r0), r = bsc * rt;
  const int ix = min((int)x, Nx - 2), iy = min((int)y, Ny - 2);
  P.fx = (T)(x - ix);
  P.fy = (T)(y - iy);
  const int o00 = tile32(ix, iy, Nx);
  const int sx = (ix & 3) == 3 ? 13 : 1, sy = (iy & 3) == 3 ? rowstep : 4;
  const int o01 = o00 + sx, o10 = o00 + sy, o11 = o10 + sx;
  if constexpr (ONE) {  // first-order Taylor step about the only node
    P.dr = (T)fma(bsc, rt, -nv.rlo);
    const char* bb = reinterpret_cast<const char*>(nv.node(0));
    auto at = [&](int o) -> const cpx<T>* {
      return reinterpret_cast<const cpx<T>*>(bb + (uint32_t)o * (uint32_t)sizeof(cpx<T>));
    };
    if constexpr (kF32) {
      // the left and right taps of a row are neighbours in the tile unless
      // ix % 4 == 3: one 16-byte load per row, and a second tap load only on
      // lanes whose cell crosses a tile column (the 16-byte load stays
      // inside the map: ix + 1 <= Nx - 1).  The crossing taps go to
      // registers of their own, so their loads need not wait for the
      // 16-byte loads to land.
      P.top = *reinterpret_cast<const f4a8*>(at(o00));
      P.bot = *reinterpret_cast<const f4a8*>(at(o10));
      P.cross = sx != 1;
      P.c[1] = cpx<T>{};
      P.c[3] = cpx<T>{};
      if (P.cross) {
        P.c[1] = *at(o01);
        P.c[3] = *at(o11);
      }
    } else {
      P.c[0] = *at(o00); P.c[1] = *at(o01); P.c[2] = *at(o10); P.c[3] = *at(o11);
    }
  } else {
    // node interval k, k+1; Taylor about the end node outside [rlo, rhi].
    // A NaN radius counts as below the first node (its Taylor step gives a
    // NaN mu, which flags the curve), and a Taylor step reads node ka twice,
    // so kb never passes the last node, also when nn == 1.
    const int nn = nv.nn;
    int k = 0;
    while (k < nn - 2 && nv.rad[k + 1] <= r) ++k;
    const bool lo = !(r > nv.rlo), hi = !lo && r >= nv.rhi;
    const int ka = hi ? nn - 1 : k, kb = (lo || hi) ? ka : k + 1;
    const double ra = nv.rad[ka], h = nv.rad[kb] - ra;
    T* w = P.w;  // wa, wda, wb, wdb, za, zda, zb, zdb
    if (ka == kb) {
      const T dr = (T)(r - ra);
      w[0] = 1; w[1] = dr; w[2] = 0; w[3] = 0;
      w[4] = 0; w[5] = 1; w[6] = 0; w[7] = 0;
    } else {
      // FP32 maps: 1/h from the table; FP64 maps keep the division
      const double ih = kF32 ? nv.ih[ka] : 1.0 / h;
      const T S = (T)(kF32 ? (r - ra) * ih : (r - ra) / h), S2 = S * S, S3 = S2 * S;
      const T H = (T)h, iH = (T)ih;
      w[0] = 2 * S3 - 3 * S2 + 1; w[1] = (S3 - 2 * S2 + S) * H;
      w[2] = -2 * S3 + 3 * S2;    w[3] = (S3 - S2) * H;
      w[4] = (6 * S2 - 6 * S) * iH; w[5] = 3 * S2 - 4 * S + 1;
      w[6] = (-6 * S2 + 6 * S) * iH; w[7] = 3 * S2 - 2 * S;
    }
    const cpx<T>* ma = nv.node(ka);
    const cpx<T>* mb = nv.node(kb);
    const int off[4] = {o00, o01, o10, o11};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      P.c[q] = ma[off[q]];
      P.c[4 + q] = mb[off[q]];
    }
  }
  return P;
}

// The second half of a sample: adds its eight FP64 terms to s and raises
// flag when mu is not positive and finite.
template <class T, bool ONE>
__device__ __forceinline__ void lc_accum(const SampleTaps<T>& P, const CurveRec& R, const LnEntry* lnt,
                                         double (&s)[8], bool& flag) {
#pragma clang fp contract(off)
  constexpr bool kF32 = sizeof(T) == 4;
  constexpr double kMag = -1.08573620475812959810;  // -2.5 / ln 10
  constexpr double kDc = 2.17147240951625919620;    //  5 / ln 10
  const T fx = P.fx, fy = P.fy, gx = (T)1 - fx, gy = (T)1 - fy;
  // corner values u (mu) and v (dmu/dr) from one or two radius nodes
  T u[4], v[4];
  if constexpr (ONE) {
    const T dr = P.dr;
    cpx<T> c0, c1, c2, c3;
    if constexpr (kF32) {
      c0 = cpx<T>{P.top.x, P.top.y};
      c2 = cpx<T>{P.bot.x, P.bot.y};
      c1 = P.cross ? P.c[1] : cpx<T>{P.top.z, P.top.w};
      c3 = P.cross ? P.c[3] : cpx<T>{P.bot.z, P.bot.w};
    } else {
      c0 = P.c[0]; c1 = P.c[1]; c2 = P.c[2]; c3 = P.c[3];
    }
    u[0] = fma(dr, c0.y, c0.x); v[0] = c0.y;
    u[1] = fma(dr, c1.y, c1.x); v[1] = c1.y;
    u[2] = fma(dr, c2.y, c2.x); v[2] = c2.y;
    u[3] = fma(dr, c3.y, c3.x); v[3] = c3.y;
  } else {
    const T* w = P.w;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const cpx<T> A = P.c[q], Bq = P.c[4 + q];
      u[q] = fma(w[3], Bq.y, fma(w[2], Bq.x, fma(w[1], A.y, w[0] * A.x)));
      v[q] = fma(w[7], Bq.y, fma(w[6], Bq.x, fma(w[5], A.y, w[4] * A.x)));
    }
  }
  // bilinear value and slopes: gy (gx u0 + fx u1) + fy (gx u2 + fx u3), ...
  const T mu_t = fma(gy, fma(gx, u[0], fx * u[1]), fy * fma(gx, u[2], fx * u[3]));
  const T mx_t = fma(gy, u[1] - u[0], fy * (u[3] - u[2]));
  const T my_t = fma(gx, u[2] - u[0], fx * (u[3] - u[1]));
  const T mr_t = fma(gy, fma(gx, v[0], fx * v[1]), fy * fma(gx, v[2], fx * v[3]));
  flag |= !(mu_t > (T)0 && mu_t < (T)INFINITY);
  double lnmu, rmu;
  if constexpr (kF32) {
    lnmu = ln_f32(mu_t, lnt);
    rmu = (double)__builtin_amdgcn_rcpf(mu_t);
  } else {
    lnmu = ln_pos(mu_t);
    rmu = 1.0 / mu_t;
  }
  const double te = P.te, eis = P.eis;
  const double rho = (P.em - fma(kMag, lnmu, R.dm)) * eis;
  const double q = rho * eis;
  const double dc = kDc * q * rmu, tdc = dc * te;
  const double mx = (double)mx_t, my = (double)my_t, mr = (double)mr_t;
  s[0] = fma(rho, rho, s[0]);
  s[1] = fma(dc, mx, s[1]);
  s[2] = fma(dc, my, s[2]);
  s[3] = fma(tdc, mx, s[3]);
  s[4] = fma(tdc, my, s[4]);
  s[5] = fma(dc, mr, s[5]);
  s[6] = fma(tdc, mr, s[6]);
  s[7] += q;
}

// The band-b epochs [e0, e1) of one curve group, strided over its LANES
// lanes: from the wavefront's LDS copy `wsep` of one (observation set,
// band) segment's records when every curve of the wavefront uses the staged
// segment st = o * B + b (the copy is refreshed first when the wavefront's
// groups agree on a set whose band-b segment has at most kStageEpochs
// epochs), otherwise from global memory one record ahead.  st is uniform
// per wavefront.
// PIPE: two epochs in flight per lane (the taps of the next are loaded while
// this one's sums are formed; P and Q alternate, so no copies).  It pays
// with several nodes (eight tap loads a sample) and on FP64 maps, 6-8% a
// launch; the FP32 single-node kernel is as fast without it at 16 lanes and
// slower at 32 (128 VGPRs).  The results are the same either way.
template <class T, bool ONE, int LANES, bool PIPE = !(ONE && sizeof(T) == 4), class NodeAt>
__device__ __forceinline__ void lc_band_epochs(const NodeView<NodeAt>& nv, int Nx, int Ny, double bsc,
                                               const CurveRec& R, bool act, bool ok, int e0, int e1, int b,
                                               int B, int nobs, const int* __restrict__ obs_off, int nepoch,
                                               const Epoch* __restrict__ ep, Epoch* wsep, int& st_o,
                                               const LnEntry* lnt, double (&s)[8], bool& flag) {
  const int lane = threadIdx.x & 63, li = lane % LANES;
  const int rowstep = 4 * Nx - 12;
  const int cand = __builtin_amdgcn_readfirstlane(act ? R.o : -1);
  if (cand >= 0 && cand < nobs && cand * B + b != st_o && __all(!act || R.o == cand)) {
    const int c0 = obs_off[(size_t)cand * B + b], c1 = obs_off[(size_t)cand * B + b + 1];
    if (c0 >= 0 && c1 <= nepoch && c0 <= c1 && c1 - c0 <= kStageEpochs) {
      for (int q = lane; q < c1 - c0; q += 64) wsep[q] = ep[c0 + q];
      st_o = cand * B + b;
    }
  }
  __builtin_amdgcn_wave_barrier();
  const int jend = ok ? e1 : e0;
  int j = e0 + li;
  if (__all(!ok || R.o * B + b == st_o)) {
    if constexpr (!PIPE) {
      for (; j < jend; j += LANES) {
        const Epoch E = wsep[j - e0];
        lc_accum<T, ONE>(lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, E.t, E.m, E.is), R, lnt, s, flag);
      }
    } else if (j < jend) {
      SampleTaps<T> P, Q;
      {
        const Epoch E = wsep[j - e0];
        P = lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, E.t, E.m, E.is);
      }
      while (true) {
        if (j + LANES < jend) {
          const Epoch E = wsep[j + LANES - e0];
          Q = lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, E.t, E.m, E.is);
        }
        lc_accum<T, ONE>(P, R, lnt, s, flag);
        j += LANES;
        if (j >= jend) break;
        if (j + LANES < jend) {
          const Epoch E = wsep[j + LANES - e0];
          P = lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, E.t, E.m, E.is);
        }
        lc_accum<T, ONE>(Q, R, lnt, s, flag);
        j += LANES;
        if (j >= jend) break;
      }
    }
  } else if constexpr (!PIPE) {
    Epoch en = j < jend ? ep[j] : Epoch{0.0, 0.0, 0.0, 0.0};
    for (; j < jend; j += LANES) {
      const Epoch E = en;
      if (j + LANES < jend) en = ep[j + LANES];
      lc_accum<T, ONE>(lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, E.t, E.m, E.is), R, lnt, s, flag);
    }
  } else if (j < jend) {
    // the record of the epoch after next is read one sample ahead of its
    // tap loads
    SampleTaps<T> P, Q;
    Epoch en;
    {
      const Epoch E = ep[j];
      P = lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, E.t, E.m, E.is);
      if (j + LANES < jend) en = ep[j + LANES];
    }
    while (true) {
      if (j + LANES < jend) {
        Q = lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, en.t, en.m, en.is);
        if (j + 2 * LANES < jend) en = ep[j + 2 * LANES];
      }
      lc_accum<T, ONE>(P, R, lnt, s, flag);
      j += LANES;
      if (j >= jend) break;
      if (j + LANES < jend) {
        P = lc_fetch<T, ONE>(nv, Nx, Ny, rowstep, bsc, R, en.t, en.m, en.is);
        if (j + 2 * LANES < jend) en = ep[j + 2 * LANES];
      }
      lc_accum<T, ONE>(Q, R, lnt, s, flag);
      j += LANES;
      if (j >= jend) break;
    }
  }
}

// Band-b epoch range of curve R and whether it lies in the pixel-centre hull
// of an Nx x Ny map at the band's earliest and latest epoch (its position
// is monotone in t, so every epoch in between lies inside as well).
__device__ __forceinline__ bool lc_band_range(const CurveRec& R, int b, int B, int nobs,
                                              const int* __restrict__ obs_off, int nepoch,
                                              const double2* __restrict__ rng, int Nx, int Ny, int& e0,
                                              int& e1) {
  const int o = R.o;
  const bool okobs = o >= 0 && o < nobs;
  e0 = okobs ? obs_off[(size_t)o * B + b] : 0;
  e1 = okobs ? obs_off[(size_t)o * B + b + 1] : 0;
  bool ok = okobs && e0 >= 0 && e1 <= nepoch && e0 <= e1;
  if (ok && e0 < e1) {
    const double xmax = (double)(Nx - 1), ymax = (double)(Ny - 1);
    const double2 tr = rng[(size_t)o * B + b];
    const double xa = fma(R.vct, tr.x, R.x0), xb = fma(R.vct, tr.y, R.x0);
    const double ya = fma(R.vst, tr.x, R.y0), yb = fma(R.vst, tr.y, R.y0);
    ok = xa >= 0.0 && xa <= xmax && xb >= 0.0 && xb <= xmax && ya >= 0.0 && ya <= ymax && yb >= 0.0 &&
         yb <= ymax;
  }
  return ok;
}

// One (profile p, band b) launch.  A group of LANES lanes per curve record
// (64/LANES curves per wavefront, grid-stride over the records of bucket
// p), the band-b epochs strided across the group.  A curve outside the
// pixel-centre hull at the band's earliest or latest epoch is flagged and
// skips the band.  Per epoch (lc_sample):
//   mu, dmu/dx, dmu/dy, dmu/dr  from the bilinear taps of the nodes (T)
//   rho = (mobs - dm + 2.5 log10 mu) / sig,  q = rho / sig        (FP64)
//   dc = d chi2 / d mu = 5 q / (ln 10 mu)
// and the group sums, in FP64,
//   [0] rho^2  [1] dc mx  [2] dc my  [3] dc t mx  [4] dc t my
//   [5] bsc dc mr  [6] bsc dc t mr  [7] -2 q
// go to the record's accumulator (band 0 writes, later bands add).  ONE: the
// (profile, band) has a single radius node (first-order Taylor step in r).
template <class T>
struct ConsecutiveNodes {
  const cpx<T>* base;
  size_t mapsz;
  __device__ __forceinline__ const cpx<T>* operator()(int k) const { return base + mapsz * k; }
};

template <class T, bool FIRST, bool ONE, int LANES>
__global__ __launch_bounds__(256) void k_lc_band(
    const cpx<T>* __restrict__ base, int Nx, int Ny, BandNodes bn, int p, int b, int B,
    const CurveRec* __restrict__ rec, const int* __restrict__ start, const int* __restrict__ cnt,
    int nobs, const int* __restrict__ obs_off, int nepoch, const Epoch* __restrict__ ep,
    const double2* __restrict__ rng, double* __restrict__ acc, int* __restrict__ bad) {
  constexpr int GPW = 64 / LANES;
  constexpr bool kF32 = sizeof(T) == 4;
  __shared__ double srad[ONE ? 1 : kMaxBandNodes], sih[ONE ? 1 : kMaxBandNodes];
  __shared__ LnEntry lnt[kF32 ? kLnTab : 1];
  // per wavefront: the band's epoch records of one observation set
  __shared__ Epoch sep[4][kStageEpochs];
  if constexpr (!ONE) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < kMaxBandNodes; ++q) {
        srad[q] = bn.rad[q];
        sih[q] = q + 1 < bn.cnt ? 1.0 / (bn.rad[q + 1] - bn.rad[q]) : 0.0;
      }
    }
  }
  if constexpr (kF32) {
    for (int q = threadIdx.x; q < kLnTab; q += 256) lnt[q] = ln_entry(q);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, gi = lane / LANES, li = lane % LANES;
  const int pos0 = start[p], ncur = cnt[p];
  const double bsc = bn.bsc;
  const NodeView<ConsecutiveNodes<T>> nv{bn.cnt, bn.rad[0], ONE ? bn.rad[0] : srad[bn.cnt - 1],
                                         ONE ? nullptr : srad, ONE ? nullptr : sih,
                                         ConsecutiveNodes<T>{base, (size_t)Nx * Ny}};
  // Workgroups are dispatched round-robin over the 8 XCDs, so blockIdx & 7
  // names the XCD.  Each XCD owns one contiguous eighth of the (map-ordered)
  // record groups and its wavefronts stride through that range together:
  // the groups in flight on one XCD at any moment are neighbours on the map
  // and share its L2.
  const int nit = (ncur + GPW - 1) / GPW;
  const int xcd = blockIdx.x & 7, nwx = (gridDim.x >> 3) * 4;
  const int wx = (blockIdx.x >> 3) * 4 + (threadIdx.x >> 6);
  const int g0 = (int)((int64_t)nit * xcd / 8), g1 = (int)((int64_t)nit * (xcd + 1) / 8);
  Epoch* const wsep = sep[threadIdx.x >> 6];
  int st_o = -1;  // uniform per wave
  for (int w = g0 + wx; w < g1; w += nwx) {  // uniform per wave
    const int i = w * GPW + gi;
    const bool act = i < ncur;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool flag = false;
    CurveRec R{};
    int e0 = 0, e1 = 0;
    bool ok = false;
    if (act) {
      R = rec[pos0 + i];
      ok = lc_band_range(R, b, B, nobs, obs_off, nepoch, rng, Nx, Ny, e0, e1);
      flag = !ok;
    }
    lc_band_epochs<T, ONE, LANES>(nv, Nx, Ny, bsc, R, act, ok, e0, e1, b, B, nobs, obs_off, nepoch, ep, wsep,
                                  st_o, lnt, s, flag);
    const double val = group_sum8<LANES>(s, li);
    const unsigned long long bm = __ballot(flag);
    constexpr unsigned long long gmask = LANES == 64 ? ~0ull : ((1ull << (LANES & 63)) - 1);
    const bool gbad = ((bm >> (gi * LANES)) & gmask) != 0;
    if (act && li % (LANES / 8) == 0) {
      const int vi = li / (LANES / 8);
      const double sc = vi == 7 ? -2.0 : (vi == 5 || vi == 6) ? bsc : 1.0;
      double* a = acc + (size_t)(pos0 + i) * 8 + vi;
      // an explicit fma: the batched kernel forms the same sum in registers
      // or in memory, and contraction must not differ between them
      if (FIRST) *a = sc * val;
      else *a = fma(sc, val, *a);
      if (vi == 0) {
        if (FIRST) bad[pos0 + i] = gbad;
        else bad[pos0 + i] |= gbad;
      }
    }
  }
}

// ---- batched path: curves on several maps, all bands in one persistent launch
// One (map, profile, band) of a batched call: its radius nodes (device
// addresses of the node maps, radii, reciprocal spacings), the band scale
// and the map size.  one: a single node on a map below 4 GiB.
struct alignas(16) BatchJob {
  const void* node[kMaxBandNodes];
  double rad[kMaxBandNodes];
  double ih[kMaxBandNodes];
  double bsc;
  int cnt, one, Nx, Ny;
};

template <class T>
struct JobNodes {
  const BatchJob* J;
  __device__ __forceinline__ const cpx<T>* operator()(int k) const {
    return reinterpret_cast<const cpx<T>*>(J->node[k]);
  }
};

// Work queue of the persistent band kernel: per XCD a counter over its
// eighth of the record groups, handed out kBatchChunk groups at a time; a
// wavefront whose XCD's range is used up takes chunks from the other XCDs'
// counters in turn.  Counters only grow, so every wavefront stops after at
// most one failed take per counter.
constexpr int kBatchChunk = 4;

// Record groups of the batched kernel: the records of (map, profile) mp,
// [start[mp], start[mp] + cnt[mp]), form ceil(cnt[mp] / gpw) groups of gpw
// records that never straddle two (map, profile) ranges; gstart[mp] is the
// first group of mp and gstart[nmp] the number of groups (one workgroup).
// It also zeroes the 8 work-queue counters of the batched kernel.
__global__ __launch_bounds__(1024) void k_batch_groups(const int* __restrict__ cnt, int nmp, int gpw,
                                                      int* __restrict__ gstart, int* __restrict__ qctr) {
  __shared__ int s[1024];
  const int t = threadIdx.x, seg = (nmp + 1023) / 1024;
  if (t < 8) qctr[t] = 0;
  const int a = min(nmp, t * seg), e = min(nmp, a + seg);
  int sum = 0;
  for (int i = a; i < e; ++i) sum += (cnt[i] + gpw - 1) / gpw;
  s[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = s[t] - sum;
  for (int i = a; i < e; ++i) {
    gstart[i] = run;
    run += (cnt[i] + gpw - 1) / gpw;
  }
  if (t == 1023) gstart[nmp] = s[1023];
}

// All bands of every (map, profile) in one persistent launch.  Records are
// in bucket order of (map, profile, map cell); a wavefront takes chunks of
// kBatchChunk record groups from the work queue, and for each group (all of
// one (map, profile) mp, so the job tables are uniform per wavefront) runs
// the bands in order, each exactly as k_lc_band does it (the same lanes per
// curve give the same per-band group sums; band 0 sets the FP64
// accumulator, later bands add), keeping the accumulator in a register and
// writing it once.  qctr: 8 counters, zero at launch.
// PER_BAND: only band bsel, added to (bsel == 0: written into) the
// accumulator in memory, as one launch per band does it.
template <class T, int LANES, bool PER_BAND>
__global__ __launch_bounds__(256) void k_lc_batch(
    const BatchJob* __restrict__ jobs, int B, const CurveRec* __restrict__ rec, const int* __restrict__ start,
    const int* __restrict__ cnt, const int* __restrict__ gstart, int nmp, int nobs,
    const int* __restrict__ obs_off, int nepoch, const Epoch* __restrict__ ep, const double2* __restrict__ rng,
    int* __restrict__ qctr, double* __restrict__ acc, int* __restrict__ bad, int bsel) {
  constexpr int GPW = 64 / LANES;
  constexpr bool kF32 = sizeof(T) == 4;
  // one kernel holds the single- and multi-node paths, so one register
  // count: FP32 keeps the unpipelined loops (111-121 VGPRs, four waves per
  // SIMD, against 149-159 with the pipelined multi-node loop)
  constexpr bool kBatchPipe = !kF32;
  __shared__ LnEntry lnt[kF32 ? kLnTab : 1];
  __shared__ Epoch sep[4][kStageEpochs];
  if constexpr (kF32) {
    for (int q = threadIdx.x; q < kLnTab; q += 256) lnt[q] = ln_entry(q);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, gi = lane / LANES, li = lane % LANES;
  const int ngroups = gstart[nmp];
  const int nchunks = (ngroups + kBatchChunk - 1) / kBatchChunk;
  const int xcd = blockIdx.x & 7;
  Epoch* const wsep = sep[threadIdx.x >> 6];
  int st_o = -1;  // uniform per wave
  int mp = 0;     // uniform per wave; groups are taken in increasing order within a chunk
  for (int hop = 0; hop < 8; ++hop) {  // own XCD's range first, then the others'
    const int x = (xcd + hop) & 7;
    const int c0 = (int)((int64_t)nchunks * x / 8), c1 = (int)((int64_t)nchunks * (x + 1) / 8);
    while (true) {
      int ch = 0;
      if (lane == 0) ch = atomicAdd(&qctr[x], 1);
      ch = __builtin_amdgcn_readfirstlane(__shfl(ch, 0, 64)) + c0;
      if (ch >= c1) break;  // uniform per wave
      const int gend = min(ngroups, (ch + 1) * kBatchChunk);
      for (int w = ch * kBatchChunk; w < gend; ++w) {
        // the (map, profile) of group w: the last mp with gstart[mp] <= w
        int lo = 0, hi = nmp - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (gstart[mid] <= w) lo = mid;
          else hi = mid - 1;
        }
        mp = __builtin_amdgcn_readfirstlane(lo);
        const int k = (w - gstart[mp]) * GPW + gi;
        const bool act = k < cnt[mp];
        const int i = start[mp] + k;
        CurveRec R{};
        if (act) R = rec[i];
        double a8 = 0.0;  // accumulator entry vi of this lane's curve
        double psc = 0.0, pval = 0.0;  // PER_BAND: the band's term sc * val
        bool gbad = false;
        const int b0 = PER_BAND ? bsel : 0, b1 = PER_BAND ? bsel + 1 : B;
        for (int b = b0; b < b1; ++b) {
          const BatchJob* J = jobs + (size_t)mp * B + b;
          const int Nx = J->Nx, Ny = J->Ny, cntn = J->cnt;
          const double bsc = J->bsc;
          double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          bool flag = false;
          int e0 = 0, e1 = 0;
          bool ok = false;
          if (act) {
            ok = lc_band_range(R, b, B, nobs, obs_off, nepoch, rng, Nx, Ny, e0, e1);
            flag = !ok;
          }
          const NodeView<JobNodes<T>> nv{cntn, J->rad[0], J->rad[cntn - 1], J->rad, J->ih, JobNodes<T>{J}};
          if (J->one)
            lc_band_epochs<T, true, LANES, kBatchPipe>(nv, Nx, Ny, bsc, R, act, ok, e0, e1, b, B, nobs, obs_off, nepoch,
                                           ep, wsep, st_o, lnt, s, flag);
          else
            lc_band_epochs<T, false, LANES, kBatchPipe>(nv, Nx, Ny, bsc, R, act, ok, e0, e1, b, B, nobs, obs_off,
                                            nepoch, ep, wsep, st_o, lnt, s, flag);
          const double val = group_sum8<LANES>(s, li);
          const unsigned long long bm = __ballot(flag);
          constexpr unsigned long long gmask = LANES == 64 ? ~0ull : ((1ull << (LANES & 63)) - 1);
          gbad |= ((bm >> (gi * LANES)) & gmask) != 0;
          const int vi = li / (LANES / 8);
          const double sc = vi == 7 ? -2.0 : (vi == 5 || vi == 6) ? bsc : 1.0;
          if (b == 0) a8 = sc * val;
          else a8 = fma(sc, val, a8);
          psc = sc;
          pval = val;
        }
        if (act && li % (LANES / 8) == 0) {
          const int vi = li / (LANES / 8);
          if (!PER_BAND || bsel == 0) {
            acc[(size_t)i * 8 + vi] = a8;
            if (vi == 0) bad[i] = gbad;
          } else {
            acc[(size_t)i * 8 + vi] = fma(psc, pval, acc[(size_t)i * 8 + vi]);  // as k_lc_band adds it
            if (vi == 0) bad[i] |= gbad;
          }
        }
      }
    }
  }
}

// chi2 and the 7-parameter gradient of every curve (thread = curve, so the
// outputs are written in order) from the accumulator of its record:
// d/dtheta = v (cos(theta) [4] - sin(theta) [3]), d/dv = cos(theta) [3] +
// sin(theta) [4]; records of bucket P (invalid profile) and flagged ones
// give (+inf, 0).
__global__ __launch_bounds__(256) void k_lc_finish(int n, int P, const int* __restrict__ start,
                                                   const int* __restrict__ pos_of,
                                                   const double* __restrict__ params,
                                                   const double* __restrict__ acc,
                                                   const int* __restrict__ bad,
                                                   double* __restrict__ chi2, double* __restrict__ grad) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int i = pos_of[c];
  const bool ok = i < start[P] && !bad[i];
  double* g = grad + (size_t)c * 7;
  if (!ok) {
    chi2[c] = INFINITY;
#pragma unroll
    for (int q = 0; q < 7; ++q) g[q] = 0.0;
    return;
  }
  const double* a = acc + (size_t)i * 8;
  double sn, cs;
  sincos(params[(size_t)c * 7 + 2], &sn, &cs);
  const double v = params[(size_t)c * 7 + 3];
  chi2[c] = a[0];
  g[0] = a[1];
  g[1] = a[2];
  g[2] = v * (cs * a[4] - sn * a[3]);
  g[3] = cs * a[3] + sn * a[4];
  g[4] = a[5];
  g[5] = a[6];
  g[6] = a[7];
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
  MLC_REQUIRE(a.Nx >= 4 && a.Ny >= 4 && a.P >= 1 && a.B >= 1 && a.nnodes >= 1 && a.ncurves >= 0 &&
                  a.nobs >= 1 && a.nepoch >= 0,
              "lc_chi2_grad: bad sizes");
  MLC_REQUIRE(a.Nx % 4 == 0 && a.Ny % 4 == 0, "lc_chi2_grad: node maps need Nx, Ny divisible by 4");
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

void validate_fused(const Args& a) {
  validate(a);
  MLC_REQUIRE(a.P < kMaxBuckets, "lc_chi2_grad: at most 8191 profiles");
  MLC_REQUIRE((int64_t)a.Nx * a.Ny <= INT32_MAX, "lc_chi2_grad: map too large for 32-bit offsets");
  for (int i = 0; i < a.P * a.B; ++i)
    MLC_REQUIRE(a.node_cnt[i] <= kMaxBandNodes, "lc_chi2_grad: at most 16 radius nodes per (profile, band)");
}

// ---- launch-shape tuner of the band kernels
// A launch shape is the number of lanes per curve (8, 16 or 32: each curve's
// epochs are strided over that many lanes, 64/lanes curves per wavefront)
// and the grid in multiples (1 or 2) of the workgroups the CUs hold at once.
// The best shape depends on epochs per band (short bands waste lanes of wide
// groups) and map size (cache reuse), so a tuned shape is kept per key
// (kernel kind, precision, log2 epochs per curve and band, log2 map pixels).
// In MLC_TUNE_AUTO mode the first call with a new key times every candidate
// twice on the call's own inputs (HIP events on the context's stream; the
// second run counts) and keeps the fastest; later calls, and captured calls,
// reuse it.  A call being captured never tunes: without a kept shape it uses
// the default.  Every shape gives the same chi2 and gradient up to the
// order of the FP64 per-band sums (lanes per curve); one shape gives the
// same bits every time.
struct LaunchShape {
  int lanes = kDefaultLanes, gmul = 1;
  int per_band = 0;  // batched kernel: one launch per band
};
enum TuneKind { kTuneBands = MLC_TUNE_KIND_BANDS, kTuneBatch = MLC_TUNE_KIND_BATCH };

template <class F>
void with_lanes(int lanes, F&& f) {
  switch (lanes) {
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: fail(MLC_ERR_ARG, "lanes per curve must be 8, 16 or 32");
  }
}

int log2_bucket(double v) {
  if (!(v >= 1.0)) return 0;
  return std::min(255, (int)std::lround(std::log2(v)));
}

uint64_t tune_key(int kind, int precision, double epochs_per_band, double pixels) {
  return ((uint64_t)kind << 24) | ((uint64_t)(precision & 1) << 16) | ((uint64_t)log2_bucket(epochs_per_band) << 8) |
         (uint64_t)log2_bucket(pixels);
}

bool capturing(mlc_ctx* ctx) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(ctx->stream, &st), "hipStreamIsCapturing");
  return st != hipStreamCaptureStatusNone;
}

bool needs_tuning(mlc_ctx* ctx, uint64_t key) {
  return ctx->tune_mode == MLC_TUNE_AUTO && !ctx->tune_cache.count(key) && !capturing(ctx);
}

LaunchShape tuned_shape(mlc_ctx* ctx, uint64_t key) {
  LaunchShape sh;
  if (ctx->tune_mode == MLC_TUNE_FORCE) {
    sh.lanes = ctx->tune_force_lanes;
    sh.gmul = ctx->tune_force_gmul;
    sh.per_band = (ctx->tune_force_flags & MLC_TUNE_BATCH_PER_BAND) != 0;
  } else if (ctx->tune_mode == MLC_TUNE_AUTO) {
    auto it = ctx->tune_cache.find(key);
    if (it != ctx->tune_cache.end()) {
      sh.lanes = it->second.lanes;
      sh.gmul = it->second.gmul;
      sh.per_band = it->second.per_band;
    }
  }
  return sh;
}

// Times run(shape) for every candidate and keeps the fastest under `key`
// (with per_band_too, each lane and grid choice also as one launch per band).
LaunchShape tune_pick(mlc_ctx* ctx, uint64_t key, const std::function<void(LaunchShape)>& run,
                      bool per_band_too = false) {
  struct Restore {  // per-kernel timing is off while the candidates run
    mlc_ctx* c;
    bool t;
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Restore() {
      c->timing = t;
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } rs{ctx, ctx->timing};
  ctx->timing = false;
  hip_check(hipEventCreate(&rs.e[0]), "hipEventCreate");
  hip_check(hipEventCreate(&rs.e[1]), "hipEventCreate");
  LaunchShape best;
  float best_ms = INFINITY;
  for (int lanes : {8, 16, 32})
    for (int gmul : {1, 2})
      for (int pb = 0; pb <= (per_band_too ? 1 : 0); ++pb) {
      LaunchShape c;
      c.lanes = lanes;
      c.gmul = gmul;
      c.per_band = pb;
      run(c);  // first run: caches and clocks settle
      hip_check(hipEventRecord(rs.e[0], ctx->stream), "hipEventRecord");
      run(c);
      hip_check(hipEventRecord(rs.e[1], ctx->stream), "hipEventRecord");
      hip_check(hipEventSynchronize(rs.e[1]), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, rs.e[0], rs.e[1]), "hipEventElapsedTime");
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  ctx->tune_cache[key] = {best.lanes, best.gmul, best.per_band};
  return best;
}

// With `fixed` the call's scratch comes from that buffer instead of the
// context's workspace (no allocation; too small a buffer is an argument error
// raised before any launch).
//
// With `concurrent`, that work (the node maps of the combined call) runs on
// the context's stream while the curve partition, which needs none of it,
// runs beside it on the side stream (run_beside); the band launches start
// when both are done.
void run_fused(mlc_ctx* ctx, const Args& a, Workspace* fixed = nullptr,
               const std::function<void()>* concurrent = nullptr) {
  validate_fused(a);
  if (a.ncurves == 0) {
    if (concurrent) (*concurrent)();
    return;
  }
  const int P = a.P, B = a.B, n = a.ncurves;
  PartGeom pg{P, 1, 0, 0.0, 0.0, a.prof, a.params, a.obs, a.nobs, B, a.nepoch, a.obs_off, a.t};
  while (pg.L < 5 && (int64_t)P * pg.C * 4 + 1 <= kMaxBuckets) {
    ++pg.L;
    pg.C *= 4;
  }
  pg.sx = (double)(1 << pg.L) / a.Nx;
  pg.sy = (double)(1 << pg.L) / a.Ny;
  const int nblk = (int)cdiv(n, kPartChunk);
  const int nq = P * pg.C + 1;
  const int64_t m = (int64_t)nq * nblk;
  MLC_REQUIRE(m <= INT32_MAX, "lc_chi2_grad: too many curves x buckets");
  Workspace own(ctx);
  Workspace& ws = fixed ? *fixed : own;
  const int ibh = ws.add(sizeof(int) * m), ist = ws.add(sizeof(int) * 2 * (P + 1)),
            itot = ws.add(sizeof(int) * 2 * (size_t)nq),
            irec = ws.add(sizeof(CurveRec) * (size_t)n), iacc = ws.add(sizeof(double) * 8 * (size_t)n),
            ibad = ws.add(sizeof(int) * (size_t)n), iep = ws.add(sizeof(Epoch) * std::max(a.nepoch, 1)),
            irng = ws.add(sizeof(double2) * (size_t)a.nobs * B), ipos = ws.add(sizeof(int) * (size_t)n);
  ws.commit();
  int* bh = ws.get<int>(ibh);
  int *start = ws.get<int>(ist), *cnt = start + (P + 1);
  int *tot = ws.get<int>(itot), *qbase = tot + nq;
  CurveRec* rec = ws.get<CurveRec>(irec);
  double* acc = ws.get<double>(iacc);
  int* bad = ws.get<int>(ibad);
  Epoch* ep = ws.get<Epoch>(iep);
  double2* rng = ws.get<double2>(irng);
  int* pos_of = ws.get<int>(ipos);
  const size_t hl = sizeof(int) * nq;
  auto partition = [&] {
    MLC_LAUNCH(ctx, "lc_partition", k_part_count, dim3(nblk), dim3(kPartThreads), hl, pg, n, bh);
    MLC_LAUNCH(ctx, "lc_partition", k_part_colscan, dim3(cdiv(nq, kColScanQ)), dim3(kColScanQ * kColScanSeg),
               0, bh, nq, nblk, tot);
    MLC_LAUNCH(ctx, "lc_partition", k_part_base, dim3(1), dim3(1024), 0, tot, nq, P, pg.C, n, qbase, start,
               cnt);
    MLC_LAUNCH(ctx, "lc_partition", k_part_fill, dim3(nblk), dim3(kPartThreads), hl, pg, n, bh, qbase, rec,
               pos_of);
    if (a.nepoch > 0)
      MLC_LAUNCH(ctx, "lc_partition", k_epoch_prep, dim3(cdiv((int64_t)a.nobs * B, 4)), dim3(256), 0,
                 a.obs_off, a.nobs * B, a.nepoch, a.t, a.mobs, a.sig, ep, rng);
  };
  if (concurrent) run_beside(ctx, partition, *concurrent);
  else partition();
  const size_t mapsz = (size_t)a.Nx * a.Ny;
  const size_t esz = a.precision == MLC_F64 ? 2 * sizeof(double) : 2 * sizeof(float);
  const bool force_general = ctx->tune_mode == 2 && (ctx->tune_force_flags & MLC_TUNE_GENERAL_NODES);
  // every (profile, band) launch with `lanes` lanes per curve and gmul times
  // as many workgroups as the CUs hold at once (a multiple of 8, so that
  // every XCD gets the same number)
  auto launch_bands = [&](int lanes, int gmul) {
    with_lanes(lanes, [&](auto lc) {
      constexpr int L = decltype(lc)::value;
      auto grid_for = [&](const void* k) {
        const int64_t want = cdiv(n, 4 * (64 / L));
        return (int)(cdiv(std::min<int64_t>(want, (int64_t)gmul * fill_grid(ctx, resident_blocks(k, 256))), 8) *
                     8);
      };
      const std::function<void(int)> profile_bands = [&](int p) {
        for (int b = 0; b < B; ++b) {
          const int pb = p * B + b;
          BandNodes bn{};
          bn.cnt = a.node_cnt[pb];
          bn.bsc = a.band_scale[b];
          for (int k = 0; k < bn.cnt; ++k) bn.rad[k] = a.node_rad[a.node_off[pb] + k];
          const void* base = (const char*)a.nodes + esz * mapsz * a.node_off[pb];
#define MLC_BAND(TT, F, O)                                                                          \
  MLC_LAUNCH(ctx, "lc_chi2_grad", (k_lc_band<TT, F, O, L>),                                         \
             dim3(grid_for(reinterpret_cast<const void*>(&k_lc_band<TT, F, O, L>))), dim3(256), 0,   \
             (const cpx<TT>*)base, a.Nx, a.Ny, bn, p, b, B, rec, start, cnt, a.nobs, a.obs_off,      \
             a.nepoch, ep, rng, acc, bad)
#define MLC_BAND_T(TT)                                                                  \
  if (bn.cnt == 1 && !force_general && mapsz * 2 * sizeof(TT) < ((size_t)1 << 32)) {    \
    if (b == 0) MLC_BAND(TT, true, true);                                               \
    else MLC_BAND(TT, false, true);                                                     \
  } else {                                                                              \
    if (b == 0) MLC_BAND(TT, true, false);                                              \
    else MLC_BAND(TT, false, false);                                                    \
  }
          if (a.precision == MLC_F64) {
            MLC_BAND_T(double)
          } else {
            MLC_BAND_T(float)
          }
#undef MLC_BAND_T
#undef MLC_BAND
        }
      };
      // the profiles' curves and accumulators are disjoint: under graph
      // capture their band sequences become parallel branches on the spread
      // streams (4096^2, three profiles: a replay 1.1% shorter, each
      // launch's tail covered by another profile's launch).  Issued eagerly
      // the cross-stream waits cost more than that (0.9% longer), and the
      // allocation-free variant creates no streams: both keep one sequence.
      if (!fixed && P > 1 && capturing(ctx)) {
        run_spread(ctx, P, profile_bands);
      } else {
        if (!fixed && P > 1) make_spread_streams(ctx);  // for a later capture
        for (int p = 0; p < P; ++p) profile_bands(p);
      }
    });
  };
  const uint64_t key = tune_key(kTuneBands, a.precision, (double)a.nepoch / ((double)a.nobs * B), mapsz);
  LaunchShape sh;
  if (!fixed && needs_tuning(ctx, key)) {
    // (the allocation-free variant never tunes: it uses a kept shape or the
    // default)
    sh = tune_pick(ctx, key, [&](LaunchShape c) { launch_bands(c.lanes, c.gmul); });
  } else {
    sh = tuned_shape(ctx, key);
  }
  launch_bands(sh.lanes, sh.gmul);
  MLC_LAUNCH(ctx, "lc_finish", k_lc_finish, dim3(cdiv(n, 256)), dim3(256), 0, n, P, start, pos_of,
             a.params, acc, bad, a.chi2, a.grad);
}

// ---- batched call
struct BatchArgs {
  int precision;
  int nmaps;
  const int *Nx, *Ny;
  int nnodes;
  const void* const* nodes;
  const double* node_rad;
  int P, B;
  const int *node_off, *node_cnt;
  const double* band_scale;
  int ncurves;
  const double* params;
  const int *map, *prof, *obs;
  int nobs;
  const int* obs_off;
  int nepoch;
  const double *t, *mobs, *sig;
  double *chi2, *grad;
};

void run_batch(mlc_ctx* ctx, const BatchArgs& a) {
  MLC_REQUIRE(a.precision == MLC_F32 || a.precision == MLC_F64, "lc_chi2_grad_batch: bad precision");
  MLC_REQUIRE(a.nmaps >= 1 && a.P >= 1 && a.B >= 1 && a.nnodes >= 1 && a.ncurves >= 0 && a.nobs >= 1 &&
                  a.nepoch >= 0,
              "lc_chi2_grad_batch: bad sizes");
  MLC_REQUIRE(a.Nx && a.Ny && a.nodes && a.node_rad && a.node_off && a.node_cnt && a.band_scale,
              "lc_chi2_grad_batch: null table");
  MLC_REQUIRE(a.ncurves == 0 || (a.params && a.map && a.prof && a.obs && a.obs_off && a.chi2 && a.grad),
              "lc_chi2_grad_batch: null curve arrays");
  MLC_REQUIRE(a.nepoch == 0 || (a.t && a.mobs && a.sig), "lc_chi2_grad_batch: null epoch arrays");
  const int P = a.P, B = a.B, NP = a.nmaps * P, n = a.ncurves;
  MLC_REQUIRE((int64_t)NP * B <= 65536, "lc_chi2_grad_batch: at most 65536 (map, profile, band) jobs");
  size_t max_pix = 0;
  for (int m = 0; m < a.nmaps; ++m) {
    MLC_REQUIRE(a.Nx[m] >= 4 && a.Ny[m] >= 4 && a.Nx[m] % 4 == 0 && a.Ny[m] % 4 == 0,
                "lc_chi2_grad_batch: node maps need Nx, Ny >= 4 and divisible by 4");
    MLC_REQUIRE((int64_t)a.Nx[m] * a.Ny[m] <= INT32_MAX, "lc_chi2_grad_batch: map too large for 32-bit offsets");
    max_pix = std::max(max_pix, (size_t)a.Nx[m] * a.Ny[m]);
  }
  for (int k = 0; k < a.nnodes; ++k) MLC_REQUIRE(a.nodes[k], "lc_chi2_grad_batch: null node map");
  const size_t esz = a.precision == MLC_F64 ? 2 * sizeof(double) : 2 * sizeof(float);
  const bool force_general = ctx->tune_mode == MLC_TUNE_FORCE && (ctx->tune_force_flags & MLC_TUNE_GENERAL_NODES);
  // the job table; a node map serves the jobs of one map only
  std::vector<int> owner(a.nnodes, -1);
  std::vector<BatchJob> jobs((size_t)NP * B);
  for (int m = 0; m < a.nmaps; ++m)
    for (int pb = 0; pb < P * B; ++pb) {
      const int j = m * P * B + pb, b = pb % B;
      const int off = a.node_off[j], c = a.node_cnt[j];
      MLC_REQUIRE(c >= 1 && c <= kMaxBandNodes && off >= 0 && off + c <= a.nnodes,
                  "lc_chi2_grad_batch: node table entry out of range (1..16 nodes per job)");
      BatchJob J{};
      J.cnt = c;
      J.Nx = a.Nx[m];
      J.Ny = a.Ny[m];
      J.bsc = a.band_scale[b];
      J.one = c == 1 && !force_general && (size_t)J.Nx * J.Ny * esz < ((size_t)1 << 32);
      for (int k = 0; k < c; ++k) {
        MLC_REQUIRE(owner[off + k] == -1 || owner[off + k] == m,
                    "lc_chi2_grad_batch: a node map is used by jobs of two maps");
        owner[off + k] = m;
        J.node[k] = a.nodes[off + k];
        J.rad[k] = a.node_rad[off + k];
        if (k > 0)
          MLC_REQUIRE(J.rad[k] > J.rad[k - 1], "lc_chi2_grad_batch: node radii must increase");
      }
      for (int k = 0; k + 1 < c; ++k) J.ih[k] = 1.0 / (J.rad[k + 1] - J.rad[k]);
      jobs[j] = J;
    }
  std::vector<int> msz(2 * (size_t)a.nmaps);
  for (int m = 0; m < a.nmaps; ++m) {
    msz[2 * m] = a.Nx[m];
    msz[2 * m + 1] = a.Ny[m];
  }
  if (n == 0) return;
  PartGeom pg{P, 1, 0, 0.0, 0.0, a.prof, a.params, a.obs, a.nobs, B, a.nepoch, a.obs_off, a.t};
  pg.nmaps = a.nmaps;
  pg.map = a.map;
  MLC_REQUIRE(NP < kMaxBuckets, "lc_chi2_grad_batch: at most 8191 (map, profile) pairs");
  while (pg.L < 5 && (int64_t)NP * pg.C * 4 + 1 <= kMaxBuckets) {
    ++pg.L;
    pg.C *= 4;
  }
  const int nblk = (int)cdiv(n, kPartChunk);
  const int nq = NP * pg.C + 1;
  const int64_t mq = (int64_t)nq * nblk;
  MLC_REQUIRE(mq <= INT32_MAX, "lc_chi2_grad_batch: too many curves x buckets");
  const BatchJob* d_jobs =
      (const BatchJob*)const_table(ctx, kTableBatchJobs, jobs.data(), sizeof(BatchJob) * jobs.size());
  pg.msz = (const int*)const_table(ctx, kTableBatchMaps, msz.data(), sizeof(int) * msz.size());
  Workspace ws(ctx);
  const int ibh = ws.add(sizeof(int) * mq), ist = ws.add(sizeof(int) * 2 * (NP + 1)),
            itot = ws.add(sizeof(int) * 2 * (size_t)nq), irec = ws.add(sizeof(CurveRec) * (size_t)n),
            iacc = ws.add(sizeof(double) * 8 * (size_t)n), ibad = ws.add(sizeof(int) * (size_t)n),
            iep = ws.add(sizeof(Epoch) * std::max(a.nepoch, 1)),
            irng = ws.add(sizeof(double2) * (size_t)a.nobs * B), ipos = ws.add(sizeof(int) * (size_t)n),
            igs = ws.add(sizeof(int) * (NP + 1)), iq = ws.add(sizeof(int) * 8);
  ws.commit();
  int* bh = ws.get<int>(ibh);
  int *start = ws.get<int>(ist), *cnt = start + (NP + 1);
  int *tot = ws.get<int>(itot), *qbase = tot + nq;
  CurveRec* rec = ws.get<CurveRec>(irec);
  double* acc = ws.get<double>(iacc);
  int* bad = ws.get<int>(ibad);
  Epoch* ep = ws.get<Epoch>(iep);
  double2* rng = ws.get<double2>(irng);
  int* pos_of = ws.get<int>(ipos);
  int* gstart = ws.get<int>(igs);
  int* qctr = ws.get<int>(iq);
  const size_t hl = sizeof(int) * nq;
  MLC_LAUNCH(ctx, "lc_partition", k_part_count, dim3(nblk), dim3(kPartThreads), hl, pg, n, bh);
  MLC_LAUNCH(ctx, "lc_partition", k_part_colscan, dim3(cdiv(nq, kColScanQ)), dim3(kColScanQ * kColScanSeg),
             0, bh, nq, nblk, tot);
  MLC_LAUNCH(ctx, "lc_partition", k_part_base, dim3(1), dim3(1024), 0, tot, nq, NP, pg.C, n, qbase, start,
             cnt);
  MLC_LAUNCH(ctx, "lc_partition", k_part_fill, dim3(nblk), dim3(kPartThreads), hl, pg, n, bh, qbase, rec,
             pos_of);
  if (a.nepoch > 0)
    MLC_LAUNCH(ctx, "lc_partition", k_epoch_prep, dim3(cdiv((int64_t)a.nobs * B, 4)), dim3(256), 0,
               a.obs_off, a.nobs * B, a.nepoch, a.t, a.mobs, a.sig, ep, rng);
  // every band in one persistent launch, or (sh.per_band) one launch per
  // band over all maps; each launch starts from a fresh work queue
  auto launch = [&](LaunchShape sh) {
    with_lanes(sh.lanes, [&](auto lc) {
      constexpr int L = decltype(lc)::value;
      auto go = [&](auto tag, auto pbc) {
        using TT = decltype(tag);
        constexpr bool PB = decltype(pbc)::value;
        auto k = k_lc_batch<TT, L, PB>;
        const int64_t want = cdiv(n, 4 * (64 / L));
        const int grid = (int)(cdiv(std::min<int64_t>(want, (int64_t)sh.gmul *
                                                                fill_grid(ctx, resident_blocks((const void*)k, 256))),
                                    8) *
                               8);
        for (int b = 0; b < (PB ? B : 1); ++b) {
          MLC_LAUNCH(ctx, "lc_batch_groups", k_batch_groups, dim3(1), dim3(1024), 0, cnt, NP, 64 / L, gstart,
                     qctr);
          MLC_LAUNCH(ctx, "lc_batch", k, dim3(grid), dim3(256), 0, d_jobs, B, rec, start, cnt, gstart, NP, a.nobs,
                     a.obs_off, a.nepoch, ep, rng, qctr, acc, bad, b);
        }
      };
      using std::integral_constant;
      if (a.precision == MLC_F64) {
        if (sh.per_band) go(double{}, integral_constant<bool, true>{});
        else go(double{}, integral_constant<bool, false>{});
      } else {
        if (sh.per_band) go(float{}, integral_constant<bool, true>{});
        else go(float{}, integral_constant<bool, false>{});
      }
    });
  };
  const uint64_t key = tune_key(kTuneBatch, a.precision, (double)a.nepoch / ((double)a.nobs * B), (double)max_pix);
  if (needs_tuning(ctx, key)) tune_pick(ctx, key, launch, B > 1);
  launch(tuned_shape(ctx, key));
  MLC_LAUNCH(ctx, "lc_finish", k_lc_finish, dim3(cdiv(n, 256)), dim3(256), 0, n, NP, start, pos_of, a.params,
             acc, bad, a.chi2, a.grad);
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

// ---- allocation-free variant: the scratch of mlc_lc_chi2_grad_noalloc lives
// in this file-scope device array (one copy per GPU, for the life of the
// process).  Host code never uses the array's name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress, called
// once per device with that device current, and is what the kernels get.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (the context's device is current:
// guarded() made it so).
void* scratch_ptr(mlc_ctx* ctx) {
  MLC_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, "lc_chi2_grad_noalloc: device ordinal too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void*& p = g_scratch_ptr[ctx->device];
  if (!p) {
    void* q = nullptr;
    hip_check(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    MLC_REQUIRE(q != nullptr, "lc_chi2_grad_noalloc: no device address for the static scratch");
    p = q;
  }
  return p;
}

}  // namespace

namespace mlc {
ProfileParams checked_params(int shape, double rhalf, double pix, double rin, double a, int Nx, int Ny);
}

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

int mlc_lc_chi2_grad_convolve(mlc_map* map, const int* node_shape, double pix, void* d_nodes, int nnodes,
                              int P, int B, const int* node_off, const int* node_cnt,
                              const double* node_rad, const double* band_scale, int ncurves,
                              const double* d_params, const int* d_prof, const int* d_obs, int nobs,
                              const int* d_obs_off, int nepoch, const double* d_t, const double* d_mobs,
                              const double* d_sig, double* d_chi2, double* d_grad) {
  if (!map) return MLC_ERR_ARG;
  mlc_ctx* ctx = map->ctx;
  return guarded(ctx, [&] {
    Args a{map->precision, d_nodes, nnodes, map->Nx, map->Ny, P, B, node_off, node_cnt, node_rad,
           band_scale, ncurves, d_params, d_prof, d_obs, nobs, d_obs_off, nepoch, d_t, d_mobs, d_sig,
           d_chi2, d_grad};
    validate_fused(a);
    MLC_REQUIRE(node_shape, "lc_chi2_grad_convolve: null node shapes");
    std::vector<ProfileParams> Ps(nnodes);
    for (int k = 0; k < nnodes; ++k)
      Ps[k] = checked_params(node_shape[k], node_rad[k], pix, 0.0, 1.0, map->Nx, map->Ny);
    const size_t bytes = 2 * (map->precision == MLC_F64 ? sizeof(double) : sizeof(float)) *
                         (size_t)map->Nx * map->Ny;
    std::vector<void*> dst(nnodes);
    for (int k = 0; k < nnodes; ++k) dst[k] = (char*)d_nodes + bytes * k;
    // every node map first, then the curves: the node kernels run back to
    // back while the row spectrum they all read stays in the last-level
    // cache (convolving each node right before the band launch that reads
    // it let the band kernels evict the spectrum between nodes, and the step
    // took 3% longer); the curve partition runs beside the node kernels
    NodeDriver drive = [&](const std::function<void(int)>& conv, bool independent) {
      // under graph capture, independent nodes become parallel branches on
      // the spread streams (the tail of each node kernel covered by
      // another's); issued eagerly they stay in order
      const std::function<void()> nodes = [&] {
        const bool spread = independent && nnodes > 1;
        if (spread && capturing(ctx)) {
          run_spread(ctx, nnodes, conv);
        } else {
          if (spread) make_spread_streams(ctx);  // for a later capture
          for (int k = 0; k < nnodes; ++k) conv(k);
        }
      };
      run_fused(ctx, a, nullptr, &nodes);
    };
    if (map->precision == MLC_F64) fft_convolve_nodes<double>(map, Ps, dst.data(), &drive);
    else fft_convolve_nodes<float>(map, Ps, dst.data(), &drive);
  });
}

int mlc_lc_chi2_grad_noalloc(mlc_ctx* ctx, int precision, const void* d_nodes, int nnodes, int Nx, int Ny,
                             int P, int B, const int* node_off, const int* node_cnt,
                             const double* node_rad, const double* band_scale, int ncurves,
                             const double* d_params, const int* d_prof, const int* d_obs, int nobs,
                             const int* d_obs_off, int nepoch, const double* d_t, const double* d_mobs,
                             const double* d_sig, double* d_chi2, double* d_grad) {
  return guarded(ctx, [&] {
    Args a{precision, d_nodes, nnodes, Nx, Ny, P, B, node_off, node_cnt, node_rad, band_scale,
           ncurves, d_params, d_prof, d_obs, nobs, d_obs_off, nepoch, d_t, d_mobs, d_sig, d_chi2,
           d_grad};
    Workspace ws(ctx, scratch_ptr(ctx), kNoAllocScratchBytes, "lc_chi2_grad_noalloc");
    run_fused(ctx, a, &ws);
  });
}

int mlc_lc_chi2_grad_batch(mlc_ctx* ctx, int precision, int nmaps, const int* Nx, const int* Ny, int nnodes,
                           void* const* d_nodes, const double* node_rad, int P, int B, const int* node_off,
                           const int* node_cnt, const double* band_scale, int ncurves, const double* d_params,
                           const int* d_map, const int* d_prof, const int* d_obs, int nobs, const int* d_obs_off,
                           int nepoch, const double* d_t, const double* d_mobs, const double* d_sig,
                           double* d_chi2, double* d_grad) {
  return guarded(ctx, [&] {
    BatchArgs a{precision, nmaps, Nx, Ny, nnodes, (const void* const*)d_nodes, node_rad, P, B, node_off,
                node_cnt, band_scale, ncurves, d_params, d_map, d_prof, d_obs, nobs, d_obs_off, nepoch, d_t,
                d_mobs, d_sig, d_chi2, d_grad};
    run_batch(ctx, a);
  });
}

int mlc_tune_mode(mlc_ctx* ctx, int mode) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(mode == MLC_TUNE_OFF || mode == MLC_TUNE_AUTO || mode == MLC_TUNE_FORCE, "mlc_tune_mode: bad mode");
    ctx->tune_mode = mode;
  });
}

int mlc_tune_force(mlc_ctx* ctx, int lanes, int gmul, int flags) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(lanes == 8 || lanes == 16 || lanes == 32, "mlc_tune_force: lanes must be 8, 16 or 32");
    MLC_REQUIRE(gmul >= 1 && gmul <= 8, "mlc_tune_force: grid multiple must be 1..8");
    MLC_REQUIRE((flags & ~(MLC_TUNE_GENERAL_NODES | MLC_TUNE_BATCH_PER_BAND)) == 0, "mlc_tune_force: unknown flag");
    ctx->tune_force_lanes = lanes;
    ctx->tune_force_gmul = gmul;
    ctx->tune_force_flags = flags;
  });
}

int mlc_tune_lookup(mlc_ctx* ctx, int kind, int precision, double epochs_per_band, double pixels, int* lanes,
                    int* gmul, int* flags) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(lanes && gmul && flags, "mlc_tune_lookup: null output");
    MLC_REQUIRE(kind == MLC_TUNE_KIND_BANDS || kind == MLC_TUNE_KIND_BATCH, "mlc_tune_lookup: bad kind");
    auto it = ctx->tune_cache.find(tune_key(kind, precision, epochs_per_band, pixels));
    const bool hit = it != ctx->tune_cache.end();
    *lanes = hit ? it->second.lanes : 0;
    *gmul = hit ? it->second.gmul : 0;
    *flags = hit && it->second.per_band ? MLC_TUNE_BATCH_PER_BAND : 0;
  });
}

int mlc_tune_save(mlc_ctx* ctx, const char* path) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(path, "mlc_tune_save: null path");
    FILE* f = std::fopen(path, "w");
    MLC_REQUIRE(f, std::string("mlc_tune_save: cannot write ") + path);
    std::fprintf(f, "# mlc launch shapes: key lanes grid_multiple per_band\n");
    for (auto& kv : ctx->tune_cache)
      std::fprintf(f, "%llu %d %d %d\n", (unsigned long long)kv.first, kv.second.lanes, kv.second.gmul,
                   kv.second.per_band);
    const bool bad = std::fclose(f) != 0;
    MLC_REQUIRE(!bad, std::string("mlc_tune_save: write failed: ") + path);
  });
}

int mlc_tune_load(mlc_ctx* ctx, const char* path) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(path, "mlc_tune_load: null path");
    FILE* f = std::fopen(path, "r");
    MLC_REQUIRE(f, std::string("mlc_tune_load: cannot read ") + path);
    std::map<uint64_t, mlc_ctx::TunedShape> got;
    char line[256];
    bool ok = true;
    while (std::fgets(line, sizeof line, f)) {
      if (line[0] == '#' || line[0] == '\n') continue;
      unsigned long long k = 0;
      int l = 0, g = 0, pb = 0;
      const int nf = std::sscanf(line, "%llu %d %d %d", &k, &l, &g, &pb);
      if (nf < 3 || !(l == 8 || l == 16 || l == 32) || g < 1 || g > 8 || pb < 0 || pb > 1) {
        ok = false;
        break;
      }
      got[(uint64_t)k] = {l, g, pb};
    }
    std::fclose(f);
    MLC_REQUIRE(ok, std::string("mlc_tune_load: malformed line in ") + path);
    for (auto& kv : got) ctx->tune_cache[kv.first] = kv.second;
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
