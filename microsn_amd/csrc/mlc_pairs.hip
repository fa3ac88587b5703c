// mlc_pairs.hip -- the reference's pair mode and its driver-side pieces:
// expanding_source (gpu_functions.cu:82-225), kernel_z + setup_integral_GPU
// (643-658, 351-401), kernel_chi2 + calculate_chi2_GPU (700-750, 428-456),
// kernel_bin_chi2 + bin_chi2_GPU (772-795, 483-503) and getBinnedMpd
// (magnification_map_gpu.cu:45-87).
//
// Built with -ffp-contract=off: every FP64 expression is evaluated with the
// same roundings, in the same order, as the reference's CPU twins
// (cpu_functions.cpp:166-283), so the results are bit-identical to them.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mlc_internal.hpp"
#include "mlc_profile.hpp"

using namespace mlc;

namespace mlc {
void sample_points(mlc_ctx* ctx, int precision, const void* conv, int Nx, int Ny, int k, int Nprof,
                   int Nloc, const int* lx, const int* ly, double* LC, double* DLC);
void sort_f64(mlc_ctx* ctx, const double* keys, int64_t n, unsigned* perm, unsigned* rank,
              double* sorted);
void bin_edges(mlc_ctx* ctx, const double* sorted, int64_t n, const double* bins, int Nbins,
               unsigned* lower_ind, unsigned* n_per_bin);
void bin_edges_device(mlc_ctx* ctx, const double* sorted, int64_t n, const double* d_bins, int Nbins,
                      unsigned* d_upper, unsigned* d_lower, unsigned* d_npb);
void map_convolve_value(mlc_map* m, const ProfileParams& P, void* d_value);
ProfileParams checked_params(int shape, double rhalf, double pix, double rin, double a, int Nx, int Ny);
}  // namespace mlc

namespace {

// kernel_z: z[a*Nloc+b] = LCA[a*Nprof] / LCB[b*Nprof].  grid (ceil(Nloc/256), Nloc).
__global__ __launch_bounds__(256) void k_z(const double* __restrict__ LCA, const double* __restrict__ LCB,
                                           int Nloc, int Nprof, double* __restrict__ z) {
  const int b = blockIdx.x * 256 + threadIdx.x, a = blockIdx.y;
  if (b >= Nloc) return;
  z[(size_t)a * Nloc + b] = LCA[(size_t)a * Nprof] / LCB[(size_t)b * Nprof];
}

// Interpolated model magnifications (kernel_chi2 gpu_functions.cu:709-720):
//   m[i][loc] = 1 if ind[i] == -1 else LC[loc*Nprof+ind[i]] + fac[i]*DLC[loc*(Nprof-1)+ind[i]]
// stored [i][loc] so that the pair kernel reads B values coalesced.
__global__ void k_interp(const double* __restrict__ LC, const double* __restrict__ DLC, int Nloc,
                         int Nprof, int Njp, const int* __restrict__ ind,
                         const double* __restrict__ fac, double* __restrict__ m) {
  const int loc = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (loc >= Nloc) return;
  const int j = ind[i];
  double v;
  if (j == -1) v = 1;
  else v = LC[(size_t)loc * Nprof + j] + fac[i] * DLC[(size_t)loc * (Nprof - 1) + j];
  m[(size_t)i * Nloc + loc] = v;
}

// chi2[sorted[a*Nloc+b]] = sum_i ((d_i - mA_i/mB_i)/s_i)^2 (kernel_chi2
// 724-748).  One workgroup row a, threads over b.  grid (ceil(Nloc/256), Nloc).
__global__ __launch_bounds__(256) void k_chi2_pairs(const double* __restrict__ mA,
                                                    const double* __restrict__ mB, int Nloc, int Njp,
                                                    const double* __restrict__ d,
                                                    const double* __restrict__ s,
                                                    const unsigned* __restrict__ sorted,
                                                    double* __restrict__ chi2) {
  __shared__ double sa[256], sd[256], ss[256];
  const int a = blockIdx.y, b = blockIdx.x * 256 + threadIdx.x;
  for (int i = threadIdx.x; i < Njp; i += 256) {
    sa[i] = mA[(size_t)i * Nloc + a];
    sd[i] = d[i];
    ss[i] = s[i];
  }
  __syncthreads();
  if (b >= Nloc) return;
  double c = 0.0;
  for (int i = 0; i < Njp; ++i) {
    double tmp = (sd[i] - (sa[i] / mB[(size_t)i * Nloc + b])) / ss[i];
    c += tmp * tmp;
  }
  const size_t n = (size_t)Nloc * Nloc, p = (size_t)a * Nloc + b;
  const size_t dst = sorted[p];
  if (dst < n) chi2[dst] = c;  // a rank table not made by setup_integral cannot write outside
}

// Per-bin sums over contiguous chunks (<= kChunk elements, never crossing a
// bin), one workgroup per chunk; the chunk sums of a bin are added on the
// host in chunk order, so the result does not depend on scheduling.
// part[2*c] = sum exp(-0.5 chi2), part[2*c+1] = sum chi2.
constexpr int kChunk = 1 << 12;

__global__ __launch_bounds__(256) void k_bin_chunks(const double* __restrict__ chi2,
                                                    const unsigned long long* __restrict__ cstart,
                                                    const unsigned* __restrict__ clen,
                                                    double* __restrict__ part) {
  __shared__ double re[256], rc[256];
  const int c = blockIdx.x;
  const double* x = chi2 + cstart[c];
  const unsigned len = clen[c];
  double se = 0.0, sc = 0.0;
  for (unsigned i = threadIdx.x; i < len; i += 256) {
    double v = x[i];
    se += exp(-0.5 * v);
    sc += v;
  }
  re[threadIdx.x] = se;
  rc[threadIdx.x] = sc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      re[threadIdx.x] += re[threadIdx.x + w];
      rc[threadIdx.x] += rc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * c] = re[0];
    part[2 * c + 1] = rc[0];
  }
}

// The chunk table of mlc_bin_chi2 built on the device (one workgroup): bin i
// has ceil(npb[i] / kChunk) chunks, numbered consecutively in bin order;
// cfirst[i] is its first chunk, cfirst[Nbins] the number of chunks.  With
// non-decreasing edges there are at most maxc = Nbins + n / kChunk chunks;
// no entry at or past maxc is written, and the count is clamped to maxc.
__global__ __launch_bounds__(1024) void k_chunk_table(const unsigned* __restrict__ lower,
                                                      const unsigned* __restrict__ npb, int nb,
                                                      unsigned maxc,
                                                      unsigned long long* __restrict__ cstart,
                                                      unsigned* __restrict__ clen,
                                                      unsigned* __restrict__ cfirst) {
  __shared__ unsigned s[1024];
  const int t = threadIdx.x, seg = (nb + 1023) / 1024;
  const int a = min(nb, t * seg), e = min(nb, a + seg);
  unsigned sum = 0;
  for (int i = a; i < e; ++i) sum += (npb[i] + kChunk - 1) / kChunk;
  s[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const unsigned v = t >= o ? s[t - o] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  unsigned c = s[t] - sum;
  for (int i = a; i < e; ++i) {
    cfirst[i] = min(c, maxc);
    unsigned long long st = lower[i];
    for (unsigned left = npb[i]; left > 0 && c < maxc;) {
      const unsigned len = min(left, (unsigned)kChunk);
      cstart[c] = st;
      clen[c] = len;
      ++c;
      st += len;
      left -= len;
    }
  }
  if (t == 1023) cfirst[nb] = min(s[1023], maxc);
}

// k_bin_chunks over a device chunk table of at most gridDim.x chunks.
__global__ __launch_bounds__(256) void k_bin_chunks_dev(const double* __restrict__ chi2,
                                                        const unsigned long long* __restrict__ cstart,
                                                        const unsigned* __restrict__ clen,
                                                        const unsigned* __restrict__ cfirst, int nb,
                                                        double* __restrict__ part) {
  if (blockIdx.x >= cfirst[nb]) return;  // uniform per workgroup
  __shared__ double re[256], rc[256];
  const int c = blockIdx.x;
  const double* x = chi2 + cstart[c];
  const unsigned len = clen[c];
  double se = 0.0, sc = 0.0;
  for (unsigned i = threadIdx.x; i < len; i += 256) {
    double v = x[i];
    se += exp(-0.5 * v);
    sc += v;
  }
  re[threadIdx.x] = se;
  rc[threadIdx.x] = sc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      re[threadIdx.x] += re[threadIdx.x + w];
      rc[threadIdx.x] += rc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * c] = re[0];
    part[2 * c + 1] = rc[0];
  }
}

// Per-bin means from the chunk sums, added in chunk order as mlc_bin_chi2
// does on the host; an empty bin gives the reference kernel's sentinel.
__global__ void k_bin_finish(const double* __restrict__ part, const unsigned* __restrict__ cfirst,
                             const unsigned* __restrict__ npb, int nb, double* __restrict__ bchi2,
                             double* __restrict__ bexp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  double se = 0.0, sc = 0.0;
  for (unsigned c = cfirst[i]; c < cfirst[i + 1]; ++c) {
    se += part[2 * c];
    sc += part[2 * c + 1];
  }
  if (npb[i] > 0) {
    bexp[i] = se / npb[i];
    bchi2[i] = sc / npb[i];
  } else {  // kernel_bin_chi2 gpu_functions.cu:789-792
    bexp[i] = exp(-2000.0 / 2.0);
    bchi2[i] = 2000.0;
  }
}

// getBinnedMpd: counts of min <= v < max, v to the first bin whose right edge
// is >= v (equivalent to the reference's sort + lower_bound + upper_bound).
__global__ __launch_bounds__(256) void k_hist_mpd(const double* __restrict__ map, int64_t npix,
                                                  const double* __restrict__ edges, int nb,
                                                  double vmin, double vmax,
                                                  unsigned long long* __restrict__ counts) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* e = reinterpret_cast<double*>(smem);
  unsigned* h = reinterpret_cast<unsigned*>(e + nb);
  for (int i = threadIdx.x; i < nb; i += 256) {
    e[i] = edges[i];
    h[i] = 0;
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < npix; i += (int64_t)gridDim.x * 256) {
    double v = map[i];
    if (v >= vmin && v < vmax) {
      int lo = 0, hi = nb;  // first edge >= v
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (e[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      if (lo < nb) atomicAdd(&h[lo], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += 256)
    if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
}

}  // namespace

extern "C" {

int mlc_setup_integral(mlc_ctx* ctx, const double* d_LCA, const double* d_LCB, int Nloc, int Nprof,
                       const double* bins, int Nbins, uint32_t* d_sorted_ind, uint32_t* lower_ind,
                       uint32_t* n_per_bin) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(Nloc > 0 && Nprof > 0 && Nbins > 0, "mlc_setup_integral: bad sizes");
    MLC_REQUIRE(d_LCA && d_LCB && bins && d_sorted_ind && lower_ind && n_per_bin,
                "mlc_setup_integral: null pointer");
    const int64_t n = (int64_t)Nloc * Nloc;
    MLC_REQUIRE(n < ((int64_t)1 << 32), "mlc_setup_integral: Nloc^2 >= 2^32");
    // z, sorted z and the permutation live in one allocation of their own:
    // the sort below takes the context workspace for its passes
    void* own = nullptr;
    hip_check(hipMalloc(&own, (sizeof(double) * 2 + sizeof(unsigned)) * n), "hipMalloc(setup_integral)");
    double* z = (double*)own;
    double* zs = z + n;
    unsigned* perm = (unsigned*)(zs + n);
    try {
      MLC_LAUNCH(ctx, "pairs_z", k_z, dim3(cdiv(Nloc, 256), Nloc), dim3(256), 0, d_LCA, d_LCB, Nloc,
                 Nprof, z);
      sort_f64(ctx, z, n, perm, d_sorted_ind, zs);
      bin_edges(ctx, zs, n, bins, Nbins, lower_ind, n_per_bin);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(own);
      throw;
    }
    sync(ctx);
    hip_check(hipFree(own), "hipFree(setup_integral)");
  });
}

int mlc_calculate_chi2(mlc_ctx* ctx, int Njp, const int* indA, const int* indB, const double* facA,
                       const double* facB, const double* d, const double* s,
                       const uint32_t* d_sorted_ind, const double* d_LCA, const double* d_DLCA,
                       const double* d_LCB, const double* d_DLCB, int Nloc, int Nprof,
                       double* d_chi2) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(Njp >= 0 && Njp <= 256, "mlc_calculate_chi2: Njp must be in [0, 256]");
    MLC_REQUIRE(Nloc > 0 && Nprof > 0, "mlc_calculate_chi2: bad sizes");
    MLC_REQUIRE(d_sorted_ind && d_LCA && d_LCB && d_chi2, "mlc_calculate_chi2: null pointer");
    MLC_REQUIRE(Njp == 0 || (indA && indB && facA && facB && d && s),
                "mlc_calculate_chi2: null table");
    for (int i = 0; i < Njp; ++i) {
      MLC_REQUIRE(indA[i] >= -1 && indA[i] < Nprof - 1 && indB[i] >= -1 && indB[i] < Nprof - 1,
                  "mlc_calculate_chi2: interpolation index out of range");
      MLC_REQUIRE(s[i] != 0.0, "mlc_calculate_chi2: zero uncertainty");
    }
    const bool any_interp = std::any_of(indA, indA + Njp, [](int v) { return v != -1; }) ||
                            std::any_of(indB, indB + Njp, [](int v) { return v != -1; });
    MLC_REQUIRE(!any_interp || (d_DLCA && d_DLCB), "mlc_calculate_chi2: DLC arrays required");
    const int J = std::max(Njp, 1);
    Workspace ws(ctx);
    int iia = ws.add(sizeof(int) * J), iib = ws.add(sizeof(int) * J), ifa = ws.add(sizeof(double) * J),
        ifb = ws.add(sizeof(double) * J), idd = ws.add(sizeof(double) * J),
        iss = ws.add(sizeof(double) * J), ima = ws.add(sizeof(double) * J * Nloc),
        imb = ws.add(sizeof(double) * J * Nloc);
    ws.commit();
    if (Njp > 0) {
      h2d(ctx, ws.get<int>(iia), indA, sizeof(int) * Njp);
      h2d(ctx, ws.get<int>(iib), indB, sizeof(int) * Njp);
      h2d(ctx, ws.get<double>(ifa), facA, sizeof(double) * Njp);
      h2d(ctx, ws.get<double>(ifb), facB, sizeof(double) * Njp);
      h2d(ctx, ws.get<double>(idd), d, sizeof(double) * Njp);
      h2d(ctx, ws.get<double>(iss), s, sizeof(double) * Njp);
      MLC_LAUNCH(ctx, "pairs_interp", k_interp, dim3(cdiv(Nloc, 256), Njp), dim3(256), 0, d_LCA,
                 d_DLCA, Nloc, Nprof, Njp, ws.get<int>(iia), ws.get<double>(ifa), ws.get<double>(ima));
      MLC_LAUNCH(ctx, "pairs_interp", k_interp, dim3(cdiv(Nloc, 256), Njp), dim3(256), 0, d_LCB,
                 d_DLCB, Nloc, Nprof, Njp, ws.get<int>(iib), ws.get<double>(ifb), ws.get<double>(imb));
    }
    MLC_LAUNCH(ctx, "pairs_chi2", k_chi2_pairs, dim3(cdiv(Nloc, 256), Nloc), dim3(256), 0,
               ws.get<double>(ima), ws.get<double>(imb), Nloc, Njp, ws.get<double>(idd),
               ws.get<double>(iss), d_sorted_ind, d_chi2);
    sync(ctx);  // the host tables above live on this stack frame
  });
}

int mlc_bin_chi2(mlc_ctx* ctx, const double* d_chi2, int64_t n, const uint32_t* lower_ind,
                 const uint32_t* n_per_bin, int Nbins, double* binned_chi2, double* binned_exp) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(Nbins > 0 && lower_ind && n_per_bin && binned_chi2 && binned_exp,
                "mlc_bin_chi2: null table");
    for (int i = 0; i < Nbins; ++i)
      MLC_REQUIRE((int64_t)lower_ind[i] + n_per_bin[i] <= n,
                  "mlc_bin_chi2: bin range outside the chi2 array");
    std::vector<unsigned long long> cs;
    std::vector<unsigned> cl;
    std::vector<int> cbin;
    for (int i = 0; i < Nbins; ++i) {
      unsigned long long st = lower_ind[i];
      unsigned left = n_per_bin[i];
      while (left > 0) {
        unsigned len = std::min<unsigned>(left, kChunk);
        cs.push_back(st);
        cl.push_back(len);
        cbin.push_back(i);
        st += len;
        left -= len;
      }
    }
    MLC_REQUIRE(cs.empty() || d_chi2, "mlc_bin_chi2: null chi2");
    std::vector<double> part(2 * cs.size());
    if (!cs.empty()) {
      Workspace ws(ctx);
      int ics = ws.add(sizeof(unsigned long long) * cs.size()), icl = ws.add(sizeof(unsigned) * cl.size()),
          ipt = ws.add(sizeof(double) * part.size());
      ws.commit();
      h2d(ctx, ws.get<void>(ics), cs.data(), sizeof(unsigned long long) * cs.size());
      h2d(ctx, ws.get<void>(icl), cl.data(), sizeof(unsigned) * cl.size());
      MLC_LAUNCH(ctx, "bin_chi2_chunks", k_bin_chunks, dim3((unsigned)cs.size()), dim3(256), 0,
                 d_chi2, ws.get<unsigned long long>(ics), ws.get<unsigned>(icl), ws.get<double>(ipt));
      d2h(ctx, part.data(), ws.get<double>(ipt), sizeof(double) * part.size());
      sync(ctx);
    }
    // chunks of one bin are consecutive; add them in order
    std::vector<double> se(Nbins, 0.0), sc(Nbins, 0.0);
    for (size_t c = 0; c < cs.size(); ++c) {
      se[cbin[c]] += part[2 * c];
      sc[cbin[c]] += part[2 * c + 1];
    }
    for (int i = 0; i < Nbins; ++i) {
      if (n_per_bin[i] > 0) {
        binned_exp[i] = se[i] / n_per_bin[i];
        binned_chi2[i] = sc[i] / n_per_bin[i];
      } else {  // kernel_bin_chi2 gpu_functions.cu:789-792
        binned_exp[i] = std::exp(-2000.0 / 2.0);
        binned_chi2[i] = 2000.0;
      }
    }
  });
}

int mlc_bin_table_device(mlc_ctx* ctx, const double* d_sorted, int64_t n, const double* bins,
                         int Nbins, uint32_t* d_lower, uint32_t* d_npb, double* d_binned_chi2,
                         double* d_binned_exp) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(Nbins > 0 && bins && d_lower && d_npb && d_binned_chi2 && d_binned_exp,
                "mlc_bin_table_device: bad bins");
    MLC_REQUIRE(n >= 0 && n < ((int64_t)1 << 32) && (n == 0 || d_sorted), "mlc_bin_table_device: bad keys");
    // non-decreasing edges keep every bin count >= 0 and their sum <= n
    for (int i = 1; i < Nbins; ++i)
      MLC_REQUIRE(bins[i] >= bins[i - 1], "mlc_bin_table_device: bin edges must be non-decreasing");
    MLC_REQUIRE(bins[0] == bins[0], "mlc_bin_table_device: NaN bin edge");
    // at most one partial chunk per bin plus the full ones
    const int64_t maxc = Nbins + n / kChunk;
    Workspace ws(ctx);
    const int iu = ws.add(sizeof(unsigned) * Nbins), ics = ws.add(sizeof(unsigned long long) * maxc),
              icl = ws.add(sizeof(unsigned) * maxc), icf = ws.add(sizeof(unsigned) * (Nbins + 1)),
              ipt = ws.add(sizeof(double) * 2 * maxc);
    ws.commit();
    const double* db = static_cast<const double*>(const_table(ctx, kTableBins, bins, sizeof(double) * Nbins));
    bin_edges_device(ctx, d_sorted, n, db, Nbins, ws.get<unsigned>(iu), d_lower, d_npb);
    unsigned* cf = ws.get<unsigned>(icf);
    MLC_LAUNCH(ctx, "bin_chunk_table", k_chunk_table, dim3(1), dim3(1024), 0, d_lower, d_npb, Nbins,
               (unsigned)maxc, ws.get<unsigned long long>(ics), ws.get<unsigned>(icl), cf);
    MLC_LAUNCH(ctx, "bin_chi2_chunks", k_bin_chunks_dev, dim3((unsigned)maxc), dim3(256), 0, d_sorted,
               ws.get<unsigned long long>(ics), ws.get<unsigned>(icl), cf, Nbins, ws.get<double>(ipt));
    MLC_LAUNCH(ctx, "bin_finish", k_bin_finish, dim3(cdiv(Nbins, 256)), dim3(256), 0, ws.get<double>(ipt),
               cf, d_npb, Nbins, d_binned_chi2, d_binned_exp);
  });
}

int mlc_binned_mpd(mlc_ctx* ctx, const double* d_map, int64_t npix, int Nbins, double min,
                   double max, const double* bins, double* counts) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(npix > 0 && d_map, "mlc_binned_mpd: empty map");
    MLC_REQUIRE(Nbins > 0 && Nbins <= 4096 && bins && counts, "mlc_binned_mpd: bad bins");
    Workspace ws(ctx);
    int ie = ws.add(sizeof(double) * Nbins), ic = ws.add(sizeof(unsigned long long) * Nbins);
    ws.commit();
    h2d(ctx, ws.get<void>(ie), bins, sizeof(double) * Nbins);
    hip_check(hipMemsetAsync(ws.get<void>(ic), 0, sizeof(unsigned long long) * Nbins, ctx->stream),
              "hipMemsetAsync(counts)");
    size_t lds = (sizeof(double) + sizeof(unsigned)) * Nbins;
    int grid = std::min<int64_t>(cdiv(npix, 256), fill_grid(ctx, 8));
    MLC_LAUNCH(ctx, "mpd_hist", k_hist_mpd, dim3(grid), dim3(256), lds, d_map, npix,
               ws.get<double>(ie), Nbins, min, max, ws.get<unsigned long long>(ic));
    std::vector<unsigned long long> h(Nbins);
    d2h(ctx, h.data(), ws.get<void>(ic), sizeof(unsigned long long) * Nbins);
    sync(ctx);
    for (int i = 0; i < Nbins; ++i)
      counts[i] = (h[i] < 1) ? 1.0 / (double)npix : (double)h[i] / (double)npix;
  });
}

int mlc_expanding_source(mlc_ctx* ctx, int precision, const double* d_map, int Nx, int Ny,
                         double width, const double* sizes, int Nprof, int shape, int Nloc,
                         const int* d_loc_x, const int* d_loc_y, double* d_LC, double* d_DLC) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_expanding_source: bad precision");
    MLC_REQUIRE(d_map && sizes && Nprof >= 1 && Nloc >= 0, "mlc_expanding_source: bad arguments");
    MLC_REQUIRE(Nloc == 0 || (d_loc_x && d_loc_y && d_LC), "mlc_expanding_source: null locations");
    MLC_REQUIRE(Nprof == 1 || d_DLC, "mlc_expanding_source: null DLC");
    MLC_REQUIRE(width > 0, "mlc_expanding_source: width must be positive");
    // magnification_map.cpp:23 with Rein = 1; gpu_functions.cu:108, 184
    const double pixSizePhys = 1.0 * width / Nx;
    const double pix = mlc_parity_round(pixSizePhys);
    std::vector<ProfileParams> P;
    for (int k = 1; k < Nprof; ++k) {
      double rhalf = mlc_parity_round(sizes[k] * pixSizePhys * Nx / width);
      P.push_back(checked_params(shape, rhalf, pix, 0.0, 0.0, Nx, Ny));
    }
    sample_points(ctx, MLC_F64, d_map, Nx, Ny, 0, Nprof, Nloc, d_loc_x, d_loc_y, d_LC, d_DLC);
    if (Nprof == 1) return;
    mlc_map* m = nullptr;
    int rc = mlc_map_create(ctx, precision, d_map, Nx, Ny, precision == MLC_F32 ? 1 : 0, &m);
    if (rc != MLC_OK) fail(rc, ctx->err);
    void* conv = nullptr;
    size_t bytes = (size_t)Nx * Ny * (precision == MLC_F64 ? 8 : 4);
    hipError_t e = hipMalloc(&conv, bytes);
    if (e != hipSuccess) {
      mlc_map_destroy(m);
      hip_check(e, "hipMalloc(convolved map)");
    }
    try {
      for (int k = 1; k < Nprof; ++k) {
        map_convolve_value(m, P[k - 1], conv);
        sample_points(ctx, precision, conv, Nx, Ny, k, Nprof, Nloc, d_loc_x, d_loc_y, d_LC, d_DLC);
      }
      sync(ctx);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(conv);
      mlc_map_destroy(m);
      throw;
    }
    hip_check(hipFree(conv), "hipFree(convolved map)");
    mlc_map_destroy(m);
  });
}

}  // extern "C"
