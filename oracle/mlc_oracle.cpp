// mlc_oracle.cpp -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// A plain restatement of the reference algorithm, used by tests/ to check
// results and by bench.py as the CPU baseline.  The product library never
// links or calls this file.  Citations are relative to the reference root
// (gvernard/microSN).  OpenMP is used only over independent rows/curves, so
// results do not depend on the thread count.

#include "mlc_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

typedef std::complex<double> cd;

// ---------------------------------------------------------------------------
// parameter round trip: std::to_string (printf "%f") then std::stof
// (gpu_functions.cu:108,184; profile.hpp:144-160)
double orc_parity_round(double x) {
  std::string s = std::to_string(x);
  float f = std::stof(s);
  return (double)f;
}

static void make_even(int& n) {  // BaseProfile::makeEven profile.cpp:175-179
  if (n % 2 != 0) n += 1;
}

// Shape parameters and grid size, following createProfileFromHalfRadius
// (profile.hpp:142-166) and each shape's generateValues size rule.
int orc_profile_params(int shape, double rhalf, double pix, double extra_rin,
                       double extra_a, double* out) {
  int n = -1;
  double p0 = 0.0, p1 = 0.0;
  switch (shape) {
    case ORC_UNIFORM: {  // UniformDisc profile.cpp:190-199
      double R = std::sqrt(2.0) * rhalf;
      p0 = R;
      n = 2 * (3 + (int)std::ceil(R / pix));
      break;
    }
    case ORC_GAUSSIAN: {  // profile.cpp:233-244
      double sdev = rhalf / 1.18;
      p0 = sdev;
      double Rh = 1.18 * sdev;
      n = (int)std::ceil(4.0 * Rh / pix);
      break;
    }
    case ORC_GAUSSIAN_HOLE: {  // profile.cpp:275-288
      double Rin = extra_rin;
      double dum = (std::pow(rhalf, 2.0) - std::pow(Rin, 2.0)) / (2.0 * std::log(2.0));
      double sdev = std::sqrt(dum);
      p0 = sdev;
      p1 = Rin;
      double Rh = std::sqrt(2.0 * std::log(2.0) * std::pow(sdev, 2.0) + std::pow(Rin, 2.0));
      n = (int)std::ceil(4.0 * Rh / pix);
      break;
    }
    case ORC_THERMAL_HOLE: {  // profile.cpp:320-330
      double Rin = rhalf / 4.0;
      p0 = Rin;
      double Rh = 4.0 * Rin;
      n = (int)std::ceil(4.0 * Rh / pix);
      break;
    }
    case ORC_WAVY: {  // profile.cpp:365-375 (3 nodes, profile.hpp:156)
      double b = 3 * M_PI / (2.0 * rhalf);
      p0 = b;
      p1 = extra_a;
      double Rh = 3 * M_PI / (2.0 * b);
      n = (int)std::ceil(4.0 * Rh / pix);
      break;
    }
    case ORC_EXPONENTIAL: {  // profile.cpp:407-417
      double sigma = 0.4106 * rhalf;
      p0 = sigma;
      double Rh = sigma / 0.4106;
      n = (int)std::ceil(4.0 * Rh / pix);
      break;
    }
    default:
      return -1;
  }
  make_even(n);
  if (out) {
    out[0] = p0;
    out[1] = p1;
  }
  return n;
}

// Unnormalised profile value f and df/d(rhalf) at radius r.
static void shape_value(int shape, double r, double rhalf, const double* par,
                        double Rh_trunc, double pix_of_grid, double* f, double* df) {
  double v = 0.0, dv = 0.0;
  switch (shape) {
    case ORC_UNIFORM:  // profile.cpp:211-218.  The indicator's derivative is
                       // taken as that of a one-pixel linear edge ramp (light-curve
                       // model convention): dR/drhalf / pix inside |r - R| < pix/2.
      v = (r < par[0]) ? 1.0 : 0.0;
      if (std::fabs(r - par[0]) < 0.5 * pix_of_grid) dv = std::sqrt(2.0) / pix_of_grid;
      break;
    case ORC_GAUSSIAN: {  // profile.cpp:255-259
      double sdev = par[0];
      double s = 2 * std::pow(sdev, 2);
      v = std::exp(-r * r / s);
      dv = v * r * r / (sdev * sdev * sdev) / 1.18;
      break;
    }
    case ORC_GAUSSIAN_HOLE: {  // profile.cpp:299-307
      double sdev = par[0], Rin = par[1];
      double s = 2 * std::pow(sdev, 2);
      if (r < Rin) {
        v = 0.0;
      } else {
        v = std::exp(-r * r / s);
        double dsdev = rhalf / (2.0 * std::log(2.0) * sdev);
        dv = v * r * r / (sdev * sdev * sdev) * dsdev;
      }
      break;
    }
    case ORC_THERMAL_HOLE: {  // profile.cpp:342-350
      double Rin = par[0];
      double fac = 3.0 * Rin / (2.0 * M_PI);
      if (r > Rin) {
        v = (1.0 - std::sqrt(Rin / r)) * fac / std::pow(r, 3.0);
        double dfdRin = (-0.5 / std::sqrt(Rin * r) * fac +
                         (1.0 - std::sqrt(Rin / r)) * 3.0 / (2.0 * M_PI)) /
                        std::pow(r, 3.0);
        dv = dfdRin * 0.25;
      }
      break;
    }
    case ORC_WAVY: {  // profile.cpp:387-393
      double b = par[0], a = par[1];
      if (r < 2 * Rh_trunc) {
        v = std::pow(std::sin(b * r), 2) / (a * r);
        dv = std::sin(2.0 * b * r) / a * (-b / rhalf);
      }
      break;
    }
    case ORC_EXPONENTIAL: {  // profile.cpp:433-442 (the r_inner branch is
                             // overwritten by the next if: reference quirk)
      double sigma = par[0];
      if (r < 2 * Rh_trunc) {
        double u = std::pow(r / sigma, 0.75);
        double eu = std::exp(u);
        v = 1.0 / (eu - 1.0);
        dv = 0.75 * u * eu / (sigma * (eu - 1.0) * (eu - 1.0)) * 0.4106;
      }
      break;
    }
  }
  *f = v;
  if (df) *df = dv;
}

static double trunc_radius(int shape, const double* par) {
  switch (shape) {
    case ORC_WAVY: return 3 * M_PI / (2.0 * par[0]);
    case ORC_EXPONENTIAL: return par[0] / 0.4106;
    default: return 0.0;
  }
}

int orc_profile_values_deriv(int shape, double rhalf, double pix,
                             double extra_rin, double extra_a, double* out,
                             double* dout, int n_cap) {
  double par[2];
  int n = orc_profile_params(shape, rhalf, pix, extra_rin, extra_a, par);
  if (n < 0) return -1;
  if (n > n_cap) return -2;
  double Rh = trunc_radius(shape, par);
  // createGrid (profile.cpp:143-157): pixel centres at (j-n/2)*pix + pix/2;
  // project() with incl = orient = 0 is the identity.
  double sum = 0.0, dsum = 0.0;
  for (int i = 0; i < n; ++i) {
    double y = (i - n / 2) * pix + pix / 2.0;
    for (int j = 0; j < n; ++j) {
      double x = (j - n / 2) * pix + pix / 2.0;
      double r = std::hypot(x, y);
      double f, df;
      shape_value(shape, r, rhalf, par, Rh, pix, &f, &df);
      out[i * n + j] = f;
      if (dout) dout[i * n + j] = df;
      sum += f;
      dsum += df;
    }
  }
  // BaseProfile::normalize (profile.cpp:158-174): row-major sum, divide.
  for (int i = 0; i < n * n; ++i) {
    double K = out[i] / sum;
    if (dout) dout[i] = (dout[i] - K * dsum) / sum;
    out[i] = K;
  }
  return n;
}

int orc_profile_values(int shape, double rhalf, double pix, double extra_rin,
                       double extra_a, double* out, int n_cap) {
  return orc_profile_values_deriv(shape, rhalf, pix, extra_rin, extra_a, out,
                                  nullptr, n_cap);
}

// Kernel::setKernel (magnification_map.cpp:172-194)
void orc_set_kernel(const double* profile, int n, int Nx, int Ny, double* kernel) {
  int fNx = n, hNx = n / 2, hNy = n / 2;
  std::memset(kernel, 0, sizeof(double) * (size_t)Nx * Ny);
  for (int i = 0; i < hNy; ++i) {
    for (int j = 0; j < hNx; ++j) {
      kernel[(size_t)i * Nx + j] = profile[hNy * fNx + hNx + i * fNx + j];
      kernel[(size_t)Nx - hNx + (size_t)i * Nx + j] = profile[hNy * fNx + i * fNx + j];
      kernel[(size_t)Nx * (Ny - hNy) + (size_t)i * Nx + j] = profile[hNx + fNx * i + j];
      kernel[(size_t)Nx * (Ny - hNy) + Nx - hNx + (size_t)i * Nx + j] = profile[fNx * i + j];
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed-radix recursive Cooley-Tukey DFT (textbook decimation in time),
// roots taken from one table of the top-level length.
struct FftPlan {
  int n;
  std::vector<cd> w;  // w[k] = exp(-2 pi i k / n)
  std::vector<int> fac;
};

static FftPlan make_plan(int n) {
  FftPlan p;
  p.n = n;
  p.w.resize(n);
  for (int k = 0; k < n; ++k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    p.w[k] = cd((double)std::cos(a), (double)std::sin(a));
  }
  int m = n;
  while (m > 1) {
    int r = 0;
    if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else {
      for (int q = 3; q <= m; q += 2)
        if (m % q == 0) { r = q; break; }
    }
    p.fac.push_back(r);
    m /= r;
  }
  p.fac.push_back(1);
  return p;
}

static inline cd root(const FftPlan& p, long long e, int sign) {
  cd w = p.w[(size_t)(e % p.n)];
  return sign < 0 ? w : std::conj(w);
}

static void fft_rec(const cd* in, long is, cd* out, int n, long wstep,
                    const FftPlan& p, int sign, const int* fac) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int r = fac[0];
  int m = n / r;
  for (int j = 0; j < r; ++j)
    fft_rec(in + j * is, is * r, out + (size_t)j * m, m, wstep * r, p, sign, fac + 1);
  std::vector<cd> t(r);
  long rstep = (long)p.n / r;
  for (int k = 0; k < m; ++k) {
    for (int j = 0; j < r; ++j) t[j] = out[(size_t)j * m + k] * root(p, (long long)j * k * wstep, sign);
    for (int q = 0; q < r; ++q) {
      cd acc = 0.0;
      for (int j = 0; j < r; ++j) acc += t[j] * root(p, (long long)j * q * rstep, sign);
      out[k + (size_t)q * m] = acc;
    }
  }
}

static void fft_run(cd* data, int n, int sign, const FftPlan& p, std::vector<cd>& tmp) {
  tmp.assign(data, data + n);
  fft_rec(tmp.data(), 1, data, n, 1, p, sign, p.fac.data());
}

void orc_fft1d(double* data, int n, int sign) {
  FftPlan p = make_plan(n);
  std::vector<cd> tmp;
  fft_run(reinterpret_cast<cd*>(data), n, sign, p, tmp);
}

static void fft2d(std::vector<cd>& a, int Nx, int Ny, int sign) {
  FftPlan px = make_plan(Nx), py = make_plan(Ny);
#pragma omp parallel
  {
    std::vector<cd> tmp;
#pragma omp for schedule(static)
    for (int y = 0; y < Ny; ++y) fft_run(&a[(size_t)y * Nx], Nx, sign, px, tmp);
  }
#pragma omp parallel
  {
    std::vector<cd> col(Ny), tmp;
#pragma omp for schedule(static)
    for (int x = 0; x < Nx; ++x) {
      for (int y = 0; y < Ny; ++y) col[y] = a[(size_t)y * Nx + x];
      fft_run(col.data(), Ny, sign, py, tmp);
      for (int y = 0; y < Ny; ++y) a[(size_t)y * Nx + x] = col[y];
    }
  }
}

void orc_fft_convolve(const double* map, const double* kernel, int Nx, int Ny, double* out) {
  size_t N = (size_t)Nx * Ny;
  std::vector<cd> A(N), K(N);
  for (size_t i = 0; i < N; ++i) {
    A[i] = cd(map[i], 0.0);
    K[i] = cd(kernel[i], 0.0);
  }
  fft2d(A, Nx, Ny, -1);
  fft2d(K, Nx, Ny, -1);
  double norm = (double)Nx * Ny;
  for (size_t i = 0; i < N; ++i) K[i] = A[i] * K[i] / norm;  // kernel_multiplyFFTs
  fft2d(K, Nx, Ny, +1);
  for (size_t i = 0; i < N; ++i) out[i] = K[i].real();
}

void orc_direct_convolve(const double* map, const double* kernel, int Nx, int Ny, double* out) {
  for (int y = 0; y < Ny; ++y)
    for (int x = 0; x < Nx; ++x) {
      double acc = 0.0;
      for (int qy = 0; qy < Ny; ++qy)
        for (int qx = 0; qx < Nx; ++qx) {
          double k = kernel[(size_t)qy * Nx + qx];
          if (k == 0.0) continue;
          int sy = (y - qy + Ny) % Ny, sx = (x - qx + Nx) % Nx;
          acc += k * map[(size_t)sy * Nx + sx];
        }
      out[(size_t)y * Nx + x] = acc;
    }
}

// ---------------------------------------------------------------------------
// locations.cpp:6-14
void orc_locations_on_grid(int Nside, int Nmap, int Noff, int* x, int* y) {
  double dd = (Nmap - 2 * Noff) / Nside;  // integer division, as the reference
  for (int i = 0; i < Nside; ++i)
    for (int j = 0; j < Nside; ++j) {
      x[i * Nside + j] = Noff + ((int)std::floor(dd / 2.0 + j * dd));
      y[i * Nside + j] = Noff + ((int)std::floor(dd / 2.0 + i * dd));
    }
}

// locations.cpp:17-25 (same libstdc++ distribution: bit-exact with GCC 11)
void orc_locations_random(int Nloc, int Nmap, int Noff, int seed, int* x, int* y) {
  int Neff = Nmap - 2 * Noff;
  std::mt19937 gen(seed);
  std::uniform_int_distribution<unsigned> distrib(0, Neff);
  for (int i = 0; i < Nloc; ++i) {
    x[i] = Noff + distrib(gen);
    y[i] = Noff + distrib(gen);
  }
}

// kernel_sample gpu_functions.cu:604-620
void orc_sample(const double* conv, int Nx, int k, int Nprof, int Nloc,
                const int* loc_x, const int* loc_y, double* LC, double* DLC) {
  for (int id = 0; id < Nloc; ++id) {
    size_t index = (size_t)(unsigned)loc_y[id] * Nx + (unsigned)loc_x[id];
    double val = conv[index];
    LC[(size_t)id * Nprof + k] = val;
    if (k > 0) DLC[(size_t)id * (Nprof - 1) + k - 1] = val - LC[(size_t)id * Nprof + k - 1];
  }
}

// expanding_source gpu_functions.cu:82-225
int orc_expanding_source(const double* map, int Nx, int Ny, double width,
                         const double* sizes, int Nprof, int shape, int Nloc,
                         const int* loc_x, const int* loc_y, double* LC, double* DLC) {
  double pixSizePhys = 1.0 * width / Nx;  // magnification_map.cpp:23, Rein = 1
  double pix = orc_parity_round(pixSizePhys);
  orc_sample(map, Nx, 0, Nprof, Nloc, loc_x, loc_y, LC, DLC);
  std::vector<double> kernel((size_t)Nx * Ny), conv((size_t)Nx * Ny);
  for (int k = 1; k < Nprof; ++k) {
    double rhalf = orc_parity_round(sizes[k] * pixSizePhys * Nx / width);
    double par[2];
    int n = orc_profile_params(shape, rhalf, pix, 0.0, 0.0, par);
    if (n < 0) return -1;
    if (n > Nx || n > Ny) return -2;
    std::vector<double> prof((size_t)n * n);
    orc_profile_values(shape, rhalf, pix, 0.0, 0.0, prof.data(), n);
    orc_set_kernel(prof.data(), n, Nx, Ny, kernel.data());
    orc_fft_convolve(map, kernel.data(), Nx, Ny, conv.data());
    orc_sample(conv.data(), Nx, k, Nprof, Nloc, loc_x, loc_y, LC, DLC);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// find_indices cpu_functions.cpp:28-46: the reference's forward walk over
// ascending times.  Defined where the reference is not: the last interval is
// closed (t == x[Nprof-1] maps to Nprof-2), t[Nd] is never read, and a point
// the walk would leave unset (t < x[0], or a time smaller than its
// predecessor's interval) is located on its own: -1 if t < x[0].
void orc_find_indices(const double* t, const double* x, int Nd, int Nprof, int* ind) {
  for (int i = 0; i < Nd; ++i) ind[i] = -1;
  int i = 0, j = 0;
  while (i < Nd && j < Nprof - 1) {
    double comp = t[i];
    bool last = (j == Nprof - 2);
    if (x[j] <= comp && (comp < x[j + 1] || (last && comp == x[j + 1]))) {
      ind[i] = j;
      i++;
    } else if (comp > x[Nprof - 1]) {
      ind[i] = -1;
      i++;
    } else if (comp < x[j]) {
      // not in the reference (which would walk off the grid): restart the
      // search for this point from the first interval
      if (comp < x[0]) {
        ind[i] = -1;
        i++;
      } else {
        j = 0;
      }
    } else {
      j++;
    }
  }
}

// setup_chi2_calculation cpu_functions.cpp:80-144
int orc_setup_chi2(double M0, double M1, double V, double Dfac, int Nd,
                   const double* d, const double* t, const double* s, int Nprof,
                   const double* x, int quirk_const_term, int* indA, int* indB,
                   double* facA, double* facB, double* new_d, double* new_s,
                   double* const_term) {
  double Rfac = 23.92865 * Dfac;
  double t2xA = V * 8.64 / (Rfac * std::sqrt(M0));
  double t2xB = V * 8.64 / (Rfac * std::sqrt(M1));
  std::vector<double> tA(Nd), tB(Nd);
  for (int i = 0; i < Nd; ++i) {
    tA[i] = t2xA * t[i];
    tB[i] = t2xB * t[i];
  }
  std::vector<int> iA(Nd), iB(Nd);
  orc_find_indices(tA.data(), x, Nd, Nprof, iA.data());
  orc_find_indices(tB.data(), x, Nd, Nprof, iB.data());
  int Njp = Nd;
  for (int i = 0; i < Nd; ++i)
    if (iA[i] == -1 && iB[i] == -1) {
      Njp = i;
      break;
    }
  for (int i = 0; i < Njp; ++i) {
    indA[i] = iA[i];
    indB[i] = iB[i];
    new_d[i] = d[i];
    new_s[i] = s[i];
    int a = iA[i];
    facA[i] = (a != -1) ? (tA[i] - x[a]) / (x[a + 1] - x[a]) : 0.0;
    int b = iB[i];
    facB[i] = (b != -1) ? (tB[i] - x[b]) / (x[b + 1] - x[b]) : 0.0;
  }
  double ct = 0.0;
  if (Njp != Nd) {
    int start = quirk_const_term ? Nd - Njp : Njp;  // cpu_functions.cpp:138
    for (int i = start; i < Nd; ++i) ct += std::pow((d[i] - 1) / s[i], 2);
  }
  *const_term = ct;
  return Njp;
}

// ---------------------------------------------------------------------------
// kernel_z gpu_functions.cu:643-658 / cpu_functions.cpp:171-176
void orc_z(const double* LCA, const double* LCB, int Nloc, int Nprof, double* z) {
  for (int a = 0; a < Nloc; ++a)
    for (int b = 0; b < Nloc; ++b)
      z[(size_t)a * Nloc + b] = LCA[(size_t)a * Nprof] / LCB[(size_t)b * Nprof];
}

static inline bool key_less(double a, double b) {
  if (std::isnan(a)) return false;
  if (std::isnan(b)) return true;
  return a < b;
}

// std::stable_sort by '<' (cpu_functions.cpp:179-181), NaN last.
void orc_stable_argsort(const double* keys, int64_t n, uint32_t* perm) {
  std::iota(perm, perm + n, 0u);
  std::stable_sort(perm, perm + n, [keys](uint32_t i1, uint32_t i2) {
    return key_less(keys[i1], keys[i2]);
  });
}

// thrust::upper_bound + adjacent_difference (gpu_functions.cu:386-396).
// NaN keys (sorted last) are never counted.
void orc_bin_edges(const double* sorted, int64_t n, const double* bins, int Nbins,
                   uint32_t* lower_ind, uint32_t* n_per_bin) {
  int64_t nvalid = n;
  while (nvalid > 0 && std::isnan(sorted[nvalid - 1])) --nvalid;
  std::vector<uint32_t> upper(Nbins);
  for (int i = 0; i < Nbins; ++i)
    upper[i] = (uint32_t)(std::upper_bound(sorted, sorted + nvalid, bins[i]) - sorted);
  lower_ind[0] = 0;
  for (int i = 1; i < Nbins; ++i) lower_ind[i] = upper[i - 1];
  n_per_bin[0] = upper[0];
  for (int i = 1; i < Nbins; ++i) n_per_bin[i] = upper[i] - upper[i - 1];
}

// setup_integral_GPU gpu_functions.cu:351-401
void orc_setup_integral(const double* LCA, const double* LCB, int Nloc, int Nprof,
                        const double* bins, int Nbins, uint32_t* sorted_ind,
                        uint32_t* lower_ind, uint32_t* n_per_bin) {
  size_t n = (size_t)Nloc * Nloc;
  std::vector<double> z(n), zs(n);
  orc_z(LCA, LCB, Nloc, Nprof, z.data());
  std::vector<uint32_t> perm(n);
  orc_stable_argsort(z.data(), (int64_t)n, perm.data());
  for (size_t r = 0; r < n; ++r) {
    sorted_ind[perm[r]] = (uint32_t)r;  // second stable sort = inverse permutation
    zs[r] = z[perm[r]];
  }
  orc_bin_edges(zs.data(), (int64_t)n, bins, Nbins, lower_ind, n_per_bin);
}

// calculate_chi2_CPU cpu_functions.cpp:213-244
void orc_chi2_pairs(int Njp, const int* indA, const int* indB, const double* facA,
                    const double* facB, const double* d, const double* s,
                    const double* LCA, const double* DLCA, const double* LCB,
                    const double* DLCB, int Nloc, int Nprof,
                    const uint32_t* sorted_ind, double* chi2) {
#pragma omp parallel for schedule(static)
  for (int a = 0; a < Nloc; ++a) {
    for (int b = 0; b < Nloc; ++b) {
      double chi2_val = 0.0;
      for (int i = 0; i < Njp; ++i) {
        double mA, mB;
        if (indA[i] == -1) mA = 1;
        else mA = LCA[(size_t)a * Nprof + indA[i]] + facA[i] * DLCA[(size_t)a * (Nprof - 1) + indA[i]];
        if (indB[i] == -1) mB = 1;
        else mB = LCB[(size_t)b * Nprof + indB[i]] + facB[i] * DLCB[(size_t)b * (Nprof - 1) + indB[i]];
        double tmp = (d[i] - (mA / mB)) / s[i];
        chi2_val += tmp * tmp;
      }
      chi2[sorted_ind[(size_t)a * Nloc + b]] = chi2_val;
    }
  }
}

// kernel_bin_chi2 gpu_functions.cu:772-795 / bin_chi2_CPU cpu_functions.cpp:265-283
void orc_bin_chi2(const double* chi2, const uint32_t* lower_ind, const uint32_t* n_per_bin,
                  int Nbins, int gpu_convention, double* binned_chi2, double* binned_exp) {
  for (int i = 0; i < Nbins; ++i) {
    if (n_per_bin[i] > 0) {
      double se = 0.0, sc = 0.0;
      for (uint32_t k = 0; k < n_per_bin[i]; ++k) {
        se += std::exp(-0.5 * chi2[lower_ind[i] + k]);
        sc += chi2[lower_ind[i] + k];
      }
      binned_exp[i] = se / n_per_bin[i];
      binned_chi2[i] = sc / n_per_bin[i];
    } else if (gpu_convention) {
      binned_exp[i] = std::exp(-2000.0 / 2.0);
      binned_chi2[i] = 2000.0;
    } else {
      binned_exp[i] = 0.0;
      binned_chi2[i] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Mpd::Mpd mpd.cpp:4-21
void orc_mpd_bins(int Nbins, double min, double max, int log_scale, double* bins) {
  if (log_scale) {
    double logmin = std::log10(min), logmax = std::log10(max);
    double logdbin = (logmax - logmin) / Nbins;
    for (int i = 0; i < Nbins; ++i) bins[i] = std::pow(10, logmin + (i + 1) * logdbin);
  } else {
    double dbin = (max - min) / Nbins;
    for (int i = 0; i < Nbins; ++i) bins[i] = min + (i + 1) * dbin;
  }
}

// integration.cpp:3-9
double orc_trapezium(int N, const double* f1, const double* f2, const double* x) {
  double integral = 0.0;
  for (int i = 1; i < N; ++i) integral += (f1[i - 1] * f2[i - 1] + f1[i] * f2[i]) * (x[i] - x[i - 1]) / 2.0;
  return integral;
}

// integration.cpp:12-42
double orc_simpson(int N, const double* f1, const double* f2, const double* x) {
  double integral = 0.0;
  for (int i = 2; i < N; i = i + 2) {
    double xp = (x[i] + x[i - 2]) / 2.0;
    double numA = (xp - x[i - 1]) * (xp - x[i]);
    double numB = (xp - x[i - 2]) * (xp - x[i]);
    double numC = (xp - x[i - 2]) * (xp - x[i - 1]);
    double denA = (x[i - 2] - x[i - 1]) * (x[i - 2] - x[i]);
    double denB = (x[i - 1] - x[i - 2]) * (x[i - 1] - x[i]);
    double denC = (x[i] - x[i - 2]) * (x[i] - x[i - 1]);
    double gA = f1[i - 2] * f2[i - 2], gB = f1[i - 1] * f2[i - 1], gC = f1[i] * f2[i];
    double gP = gA * numA / denA + gB * numB / denB + gC * numC / denC;
    integral += (gA + 4 * gP + gC) * (x[i] - x[i - 2]) / 6.0;
  }
  if (N % 2 != 0) integral += (f1[N - 2] * f2[N - 2] + f1[N - 1] * f2[N - 1]) * (x[N - 1] - x[N - 2]) / 2.0;
  return integral;
}

// Mpd::integrate + normalize mpd.cpp:47-68
void orc_mpd_normalize(int N, const double* bins, double* counts) {
  double sum = 0.0;
  for (int i = 1; i < N; ++i) sum += (counts[i] + counts[i - 1]) * (bins[i] - bins[i - 1]) / 2.0;
  for (int i = 0; i < N; ++i) counts[i] /= sum;
}

// Mpd::divide_by mpd.cpp:71-107
void orc_mpd_divide_by(int NA, const double* binsA, const double* countsA, int NB,
                       const double* binsB, const double* countsB, int NR,
                       const double* binsR, double* countsR) {
  std::vector<double> terms(NB);
  for (int i = 0; i < NR; ++i) {
    double z = binsR[i];
    for (int b = 0; b < NB; ++b) {
      double muB = binsB[b];
      double prod = z * muB;
      double pA = 0.0;
      if (prod <= binsA[0]) pA = 1e-8;
      else if (prod >= binsA[NA - 1]) pA = 1e-8;
      else {
        for (int k = 0; k < NA; ++k)
          if (binsA[k] > prod) {
            pA = countsA[k - 1] + (countsA[k] - countsA[k - 1]) * (prod - binsA[k - 1]) / (binsA[k] - binsA[k - 1]);
            break;
          }
      }
      terms[b] = muB * pA * countsB[b];
    }
    double sum = 0.0;
    for (int b = 1; b < NB; ++b) sum += (terms[b - 1] + terms[b]) * (binsB[b] - binsB[b - 1]) / 2.0;
    countsR[i] = sum;
  }
}

// MagnificationMap::getBinnedMpd magnification_map_gpu.cu:45-87
void orc_binned_mpd(const double* map, int64_t npix, int Nbins, double min, double max,
                    double* bins, double* counts) {
  double logmin = std::log10(min), logmax = std::log10(max);
  double logdbin = (logmax - logmin) / Nbins;
  for (int i = 0; i < Nbins; ++i) bins[i] = std::pow(10, logmin + (i + 1) * logdbin);
  std::vector<double> data(map, map + npix);
  std::sort(data.begin(), data.end());
  int64_t lo = std::lower_bound(data.begin(), data.end(), min) - data.begin();
  int64_t hi = std::lower_bound(data.begin(), data.end(), max) - data.begin();
  std::vector<int64_t> up(Nbins);
  for (int i = 0; i < Nbins; ++i)
    up[i] = std::upper_bound(data.begin() + lo, data.begin() + hi, bins[i]) - (data.begin() + lo);
  for (int i = 0; i < Nbins; ++i) {
    int64_t c = (i == 0) ? up[0] : up[i] - up[i - 1];
    counts[i] = (c < 1) ? 1.0 / (double)npix : (double)c / (double)npix;
  }
}

// magnification_map.cpp:33-38
void orc_gerlumph_normalise(const int32_t* counts, int64_t npix, double avgmu, double avgN,
                            double k, double g, double* out) {
  double factor = std::fabs(avgmu / avgN);
  double muth = std::fabs(1.0 / (std::pow(1.0 - k, 2) - std::pow(g, 2)));
  for (int64_t i = 0; i < npix; ++i) out[i] = (counts[i] * factor / muth);
}

// ---------------------------------------------------------------------------
// chi2/chi2_functions.cuh:35-68 (GPU) and 160-190 (CPU), float arithmetic.
void orc_chi2_multiband(const float* f_obs, const float* df_obs, const float* mu1,
                        const float* mu2, int N, int F, int T, int gpu_quirk,
                        float* chi2_all, float* likelihood_f32, double* likelihood_f64) {
  float like = 0.0f;
  double like64 = 0.0;
  for (int x = 0; x < N; ++x) {
    for (int y = 0; y < N; ++y) {
      int m2 = gpu_quirk ? x : y;
      float chi2 = 0.0f;
      for (int f = 0; f < F; ++f) {
        float part = 0.0f;
        for (int t = 0; t < T; ++t) {
          float tmp = (f_obs[f * T + t] - mu1[(size_t)x * F * T + f * T + t] /
                                              mu2[(size_t)m2 * F * T + f * T + t]) /
                      df_obs[f * T + t];
          if (gpu_quirk) part += tmp * tmp / 2;
          else chi2 += tmp * tmp / 2;
        }
        if (gpu_quirk) chi2 += part;
      }
      if (chi2_all) chi2_all[(size_t)N * y + x] = chi2;
      if (!gpu_quirk) {
        like += std::exp(-chi2);
        like64 += std::exp(-(double)chi2);
      }
    }
  }
  if (gpu_quirk) {
    // host loop chi2_functions.cuh:123-130: i (x) outer, j (y) inner
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        float c = chi2_all ? chi2_all[(size_t)N * j + i] : 0.0f;
        like += std::exp(-c);
        like64 += std::exp(-(double)c);
      }
  }
  if (likelihood_f32) *likelihood_f32 = like;
  if (likelihood_f64) *likelihood_f64 = like64;
}

// ---------------------------------------------------------------------------
// light-curve model (no reference function)
double orc_bilinear(const double* map, int Nx, int Ny, double x, double y) {
  if (!(x >= 0.0 && x <= Nx - 1 && y >= 0.0 && y <= Ny - 1)) return 0.0;
  int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
  if (x0 > Nx - 2) x0 = Nx - 2;
  if (y0 > Ny - 2) y0 = Ny - 2;
  double fx = x - x0, fy = y - y0;
  const double* r0 = map + (size_t)y0 * Nx + x0;
  const double* r1 = r0 + Nx;
  return (1 - fy) * ((1 - fx) * r0[0] + fx * r0[1]) + fy * ((1 - fx) * r1[0] + fx * r1[1]);
}

struct Tap {
  double v, dvdx, dvdy;  // bilinear value of the map component and its gradient
};

// bilinear of component c (0 value, 1 d/dr) of an interleaved (value, d/dr) map
static inline void bil2(const double* m, int Nx, int x0, int y0, double fx, double fy,
                        Tap* val, Tap* der) {
  const double* a = m + 2 * ((size_t)y0 * Nx + x0);
  const double* b = a + 2 * (size_t)Nx;
  for (int c = 0; c < 2; ++c) {
    double f00 = a[c], f01 = a[2 + c], f10 = b[c], f11 = b[2 + c];
    Tap t;
    t.v = (1 - fy) * ((1 - fx) * f00 + fx * f01) + fy * ((1 - fx) * f10 + fx * f11);
    t.dvdx = (1 - fy) * (f01 - f00) + fy * (f11 - f10);
    t.dvdy = (1 - fx) * (f10 - f00) + fx * (f11 - f01);
    if (c == 0) *val = t;
    else *der = t;
  }
}

void orc_lc_chi2_grad(const double* maps, int Nx, int Ny, int B, const int* node_off,
                      const int* node_cnt, const double* node_rad, const int* node_map,
                      const double* band_scale, int ncurves, const double* params,
                      const int* prof, const int* obs, const int* obs_off, const double* t,
                      const double* mobs, const double* sig, double* chi2, double* grad) {
  const double LN10 = std::log(10.0);
  size_t mapsz = (size_t)2 * Nx * Ny;
#pragma omp parallel for schedule(dynamic, 64)
  for (int c = 0; c < ncurves; ++c) {
    const double* P = params + (size_t)c * 7;
    double x0 = P[0], y0 = P[1], th = P[2], v = P[3], r0 = P[4], w = P[5], dm = P[6];
    double ct = std::cos(th), st = std::sin(th);
    double acc = 0.0, g[7] = {0, 0, 0, 0, 0, 0, 0};
    bool bad = false;
    int p = prof[c], o = obs[c];
    for (int b = 0; b < B && !bad; ++b) {
      int noff = node_off[p * B + b], ncnt = node_cnt[p * B + b];
      for (int e = obs_off[o * B + b]; e < obs_off[o * B + b + 1]; ++e) {
        double te = t[e];
        double x = x0 + v * ct * te, y = y0 + v * st * te;
        double r = band_scale[b] * (r0 + w * te);
        if (!(x >= 0.0 && x <= Nx - 1 && y >= 0.0 && y <= Ny - 1)) {
          bad = true;
          break;
        }
        int ix = (int)std::floor(x), iy = (int)std::floor(y);
        if (ix > Nx - 2) ix = Nx - 2;
        if (iy > Ny - 2) iy = Ny - 2;
        double fx = x - ix, fy = y - iy;
        double mu, dmu_dx, dmu_dy, dmu_dr;
        int k = 0;
        bool taylor = true;
        double s = 0.0, h = 1.0;
        if (ncnt >= 2) {
          double rl = node_rad[noff], rh = node_rad[noff + ncnt - 1];
          if (r <= rl) k = 0;
          else if (r >= rh) k = ncnt - 1;
          else {
            taylor = false;
            k = 0;
            while (k < ncnt - 2 && node_rad[noff + k + 1] <= r) ++k;
            h = node_rad[noff + k + 1] - node_rad[noff + k];
            s = (r - node_rad[noff + k]) / h;
          }
        }
        if (taylor) {
          Tap M, D;
          bil2(maps + mapsz * node_map[noff + k], Nx, ix, iy, fx, fy, &M, &D);
          double dr = r - node_rad[noff + k];
          mu = M.v + D.v * dr;
          dmu_dx = M.dvdx + D.dvdx * dr;
          dmu_dy = M.dvdy + D.dvdy * dr;
          dmu_dr = D.v;
        } else {
          Tap M0, D0, M1, D1;
          bil2(maps + mapsz * node_map[noff + k], Nx, ix, iy, fx, fy, &M0, &D0);
          bil2(maps + mapsz * node_map[noff + k + 1], Nx, ix, iy, fx, fy, &M1, &D1);
          double s2 = s * s, s3 = s2 * s;
          double h00 = 2 * s3 - 3 * s2 + 1, h10 = s3 - 2 * s2 + s;
          double h01 = -2 * s3 + 3 * s2, h11 = s3 - s2;
          double d00 = 6 * s2 - 6 * s, d10 = 3 * s2 - 4 * s + 1;
          double d01 = -6 * s2 + 6 * s, d11 = 3 * s2 - 2 * s;
          mu = h00 * M0.v + h10 * h * D0.v + h01 * M1.v + h11 * h * D1.v;
          dmu_dx = h00 * M0.dvdx + h10 * h * D0.dvdx + h01 * M1.dvdx + h11 * h * D1.dvdx;
          dmu_dy = h00 * M0.dvdy + h10 * h * D0.dvdy + h01 * M1.dvdy + h11 * h * D1.dvdy;
          dmu_dr = (d00 * M0.v + d10 * h * D0.v + d01 * M1.v + d11 * h * D1.v) / h;
        }
        if (!(mu > 0.0)) {
          bad = true;
          break;
        }
        double m = -2.5 * std::log10(mu) + dm;
        double sg = sig[e];
        double rho = (mobs[e] - m) / sg;
        acc += rho * rho;
        double dc_dmu = 2.0 * rho * 2.5 / (LN10 * mu * sg);
        g[0] += dc_dmu * dmu_dx;
        g[1] += dc_dmu * dmu_dy;
        g[2] += dc_dmu * (dmu_dx * (-v * st * te) + dmu_dy * (v * ct * te));
        g[3] += dc_dmu * (dmu_dx * ct * te + dmu_dy * st * te);
        g[4] += dc_dmu * dmu_dr * band_scale[b];
        g[5] += dc_dmu * dmu_dr * band_scale[b] * te;
        g[6] += -2.0 * rho / sg;
      }
    }
    if (bad) {
      chi2[c] = INFINITY;
      for (int q = 0; q < 7; ++q) grad[(size_t)c * 7 + q] = 0.0;
    } else {
      chi2[c] = acc;
      for (int q = 0; q < 7; ++q) grad[(size_t)c * 7 + q] = g[q];
    }
  }
}
