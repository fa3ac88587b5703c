// mlc_host.cpp -- host-side steps of the reference driver that feed the
// device path: sample locations, the chi2 set-up tables, MPD bin edges and
// the trapezium integral.  Plain C++ (no device code), exported through the
// same C ABI as the kernels.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/mlc_host.h"

extern "C" {

// locations_on_grid (locations.cpp:6-14).  The spacing is an integer
// division, as in the reference, then used as a double.
int mlc_locations_on_grid(int Nside, int Nmap, int Noff, int* x, int* y) {
  if (Nside <= 0 || Noff < 0 || Nmap - 2 * Noff <= 0 || !x || !y) return 1;
  const double step = (double)((Nmap - 2 * Noff) / Nside);
  for (int row = 0; row < Nside; ++row) {
    const int iy = Noff + (int)std::floor(step / 2.0 + row * step);
    for (int col = 0; col < Nside; ++col) {
      x[row * Nside + col] = Noff + (int)std::floor(step / 2.0 + col * step);
      y[row * Nside + col] = iy;
    }
  }
  return 0;
}

// locations_random (locations.cpp:17-25): mt19937(seed) and an inclusive
// uniform integer in [0, Nmap - 2 Noff], x drawn before y for each location.
// The integer distribution is libstdc++'s (the reference's toolchain).
int mlc_locations_random(int Nloc, int Nmap, int Noff, int seed, int* x, int* y) {
  if (Nloc < 0 || (Nloc > 0 && (!x || !y)) || Noff < 0 || Nmap - 2 * Noff < 0) return 1;
  std::mt19937 gen(seed);
  std::uniform_int_distribution<unsigned> pick(0, (unsigned)(Nmap - 2 * Noff));
  for (int i = 0; i < Nloc; ++i) {
    x[i] = Noff + (int)pick(gen);
    y[i] = Noff + (int)pick(gen);
  }
  return 0;
}

// find_indices (cpu_functions.cpp:28-46): the left index j of the size
// interval x[j] <= t < x[j+1] holding each point, -1 beyond the last size.
// For ascending t inside the grid this is the reference's forward walk; the
// cases it leaves unset or reads past t[Nd-1] for are defined here: the last
// interval is closed (t == x[Nprof-1] gives Nprof-2) and t < x[0] gives -1.
int mlc_find_indices(const double* t, const double* x, int Nd, int Nprof, int* ind) {
  if (Nd < 0 || Nprof < 1 || (Nd > 0 && (!t || !x || !ind))) return 1;
  for (int i = 0; i < Nd; ++i) {
    const double v = t[i];
    ind[i] = -1;
    if (Nprof < 2 || !(v >= x[0]) || v > x[Nprof - 1]) continue;
    int j = 0;
    while (j < Nprof - 2 && !(v < x[j + 1])) ++j;
    ind[i] = j;
  }
  return 0;
}

// setup_chi2_calculation (cpu_functions.cpp:80-144).  Rfac = 23.92865 sqrt(DS DLS/DL)
// enters through Dfac = sqrt(DS DLS / DL); t2x = V 8.64/(Rfac sqrt(M)).
int mlc_setup_chi2(double MA, double MB, double V, double Dfac, int Nd, const double* d,
                   const double* t, const double* s, int Nprof, const double* x,
                   int quirk_const_term, int* indA, int* indB, double* facA, double* facB,
                   double* new_d, double* new_s, int* Njp, double* const_term) {
  if (Nd < 0 || Nprof < 1 || !Njp || !const_term) return 1;
  if (Nd > 0 && (!d || !t || !s || !x || !indA || !indB || !facA || !facB || !new_d || !new_s))
    return 1;
  const double Rfac = 23.92865 * Dfac;
  const double kA = V * 8.64 / (Rfac * std::sqrt(MA));
  const double kB = V * 8.64 / (Rfac * std::sqrt(MB));
  std::vector<double> xa(Nd), xb(Nd);
  for (int i = 0; i < Nd; ++i) {
    xa[i] = kA * t[i];
    xb[i] = kB * t[i];
  }
  std::vector<int> ia(Nd), ib(Nd);
  mlc_find_indices(xa.data(), x, Nd, Nprof, ia.data());
  mlc_find_indices(xb.data(), x, Nd, Nprof, ib.data());
  int n = Nd;  // first point where both images have left the size grid
  for (int i = 0; i < Nd; ++i)
    if (ia[i] == -1 && ib[i] == -1) {
      n = i;
      break;
    }
  for (int i = 0; i < n; ++i) {
    indA[i] = ia[i];
    indB[i] = ib[i];
    new_d[i] = d[i];
    new_s[i] = s[i];
    facA[i] = ia[i] == -1 ? 0.0 : (xa[i] - x[ia[i]]) / (x[ia[i] + 1] - x[ia[i]]);
    facB[i] = ib[i] == -1 ? 0.0 : (xb[i] - x[ib[i]]) / (x[ib[i] + 1] - x[ib[i]]);
  }
  // constant term of the points left out; the reference starts its loop at
  // Nd - Njp (cpu_functions.cpp:138), reproduced when quirk_const_term != 0
  double c = 0.0;
  if (n != Nd)
    for (int i = quirk_const_term ? Nd - n : n; i < Nd; ++i) c += std::pow((d[i] - 1) / s[i], 2);
  *Njp = n;
  *const_term = c;
  return 0;
}

// Mpd::Mpd(Nbins, min, max, type) bin right edges (mpd.cpp:4-21).
int mlc_mpd_bins(int Nbins, double min, double max, int log_scale, double* bins) {
  if (Nbins <= 0 || !bins || !(max > min)) return 1;
  if (log_scale) {
    if (!(min > 0) || !(max > 0)) return 1;
    const double lo = std::log10(min), step = (std::log10(max) - lo) / Nbins;
    for (int i = 0; i < Nbins; ++i) bins[i] = std::pow(10, lo + (i + 1) * step);
  } else {
    const double step = (max - min) / Nbins;
    for (int i = 0; i < Nbins; ++i) bins[i] = min + (i + 1) * step;
  }
  return 0;
}

// Mpd::normalize (mpd.cpp:63-68 with integrate 47-61): counts /= trapezoid area.
int mlc_mpd_normalize(int N, const double* bins, double* counts) {
  if (N <= 0 || !bins || !counts) return 1;
  double area = 0.0;
  for (int i = 1; i < N; ++i) area += (counts[i] + counts[i - 1]) * (bins[i] - bins[i - 1]) / 2.0;
  for (int i = 0; i < N; ++i) counts[i] /= area;
  return 0;
}

// Mpd::divide_by (mpd.cpp:71-107): PDF of the ratio z = mu_A/mu_B,
// P(z) = int mu_B P_A(z mu_B) P_B(mu_B) dmu_B with P_A linearly interpolated
// and 1e-8 outside its range, trapezoid rule over the B bins.
int mlc_mpd_divide_by(int NA, const double* binsA, const double* countsA, int NB,
                      const double* binsB, const double* countsB, int NR, const double* binsR,
                      double* countsR) {
  if (NA < 2 || NB < 1 || NR < 1 || !binsA || !countsA || !binsB || !countsB || !binsR || !countsR)
    return 1;
  std::vector<double> f(NB);
  for (int r = 0; r < NR; ++r) {
    const double z = binsR[r];
    for (int b = 0; b < NB; ++b) {
      const double mu = binsB[b], arg = z * mu;
      double pa = 1e-8;
      if (arg > binsA[0] && arg < binsA[NA - 1]) {
        int k = 1;
        while (!(binsA[k] > arg)) ++k;
        pa = countsA[k - 1] + (countsA[k] - countsA[k - 1]) * (arg - binsA[k - 1]) / (binsA[k] - binsA[k - 1]);
      }
      f[b] = mu * pa * countsB[b];
    }
    double sum = 0.0;
    for (int b = 1; b < NB; ++b) sum += (f[b - 1] + f[b]) * (binsB[b] - binsB[b - 1]) / 2.0;
    countsR[r] = sum;
  }
  return 0;
}

// trapezium (integration.cpp:3-9): integral of f1*f2 over x.
double mlc_trapezium(int N, const double* f1, const double* f2, const double* x) {
  double acc = 0.0;
  for (int i = 1; i < N; ++i) acc += (f1[i - 1] * f2[i - 1] + f1[i] * f2[i]) * (x[i] - x[i - 1]) / 2.0;
  return acc;
}

// ------------------------------------------------------------------ map I/O
// GERLUMPH map directory (magnification_map.cpp:7-40): mapmeta.dat holds
// "avgmu avgN / Nx / width / k g s" (whitespace separated), map.bin Nx*Nx
// native-endian int32 ray counts.  Unlike the reference, unreadable or short
// files are reported.
int mlc_gerlumph_read(const char* dir, double* avgmu, double* avgN, int* Nx, double* width,
                      double* k, double* g, double* s, int32_t* counts, int64_t cap) {
  if (!dir || !avgmu || !avgN || !Nx || !width || !k || !g || !s) return 1;
  const std::string base(dir);
  FILE* f = std::fopen((base + "/mapmeta.dat").c_str(), "r");
  if (!f) return 2;
  const int got = std::fscanf(f, "%lf %lf %d %lf %lf %lf %lf", avgmu, avgN, Nx, width, k, g, s);
  std::fclose(f);
  if (got != 7 || *Nx <= 0) return 2;
  if (!counts) return 0;
  const int64_t n = (int64_t)*Nx * *Nx;
  if (cap < n) return 1;
  FILE* b = std::fopen((base + "/map.bin").c_str(), "rb");
  if (!b) return 3;
  const size_t rd = std::fread(counts, sizeof(int32_t), (size_t)n, b);
  std::fclose(b);
  return rd == (size_t)n ? 0 : 3;
}

// Writes a map directory in the same format (synthetic maps, tests).
int mlc_gerlumph_write(const char* dir, double avgmu, double avgN, int Nx, double width, double k,
                       double g, double s, const int32_t* counts) {
  if (!dir || Nx <= 0 || !counts) return 1;
  const std::string base(dir);
  FILE* f = std::fopen((base + "/mapmeta.dat").c_str(), "w");
  if (!f) return 2;
  std::fprintf(f, "%.17g %.17g\n%d\n%.17g\n%.17g %.17g %.17g\n", avgmu, avgN, Nx, width, k, g, s);
  std::fclose(f);
  FILE* b = std::fopen((base + "/map.bin").c_str(), "wb");
  if (!b) return 3;
  const size_t n = (size_t)Nx * Nx;
  const size_t wr = std::fwrite(counts, sizeof(int32_t), n, b);
  std::fclose(b);
  return wr == n ? 0 : 3;
}

namespace {
// One 80-character FITS header card.
void fits_card(std::string& h, const std::string& key, const std::string& value) {
  char c[81];  // fixed format: strings start in column 11, numbers end in column 30
  if (!value.empty() && value[0] == '\'')
    std::snprintf(c, sizeof(c), "%-8.8s= %-20s", key.c_str(), value.c_str());
  else
    std::snprintf(c, sizeof(c), "%-8.8s= %20s", key.c_str(), value.c_str());
  std::string card(c);
  card.resize(80, ' ');
  h += card;
}
void fits_end(std::string& h) {
  std::string card("END");
  card.resize(80, ' ');
  h += card;
  h.resize((h.size() + 2879) / 2880 * 2880, ' ');
}
// Big-endian float32 data unit, zero-padded to 2880 bytes.
void fits_data(std::string& out, const std::vector<float>& v) {
  const size_t start = out.size();
  out.resize(start + v.size() * 4);
  for (size_t i = 0; i < v.size(); ++i) {
    uint32_t u;
    std::memcpy(&u, &v[i], 4);
    out[start + 4 * i + 0] = (char)(u >> 24);
    out[start + 4 * i + 1] = (char)(u >> 16);
    out[start + 4 * i + 2] = (char)(u >> 8);
    out[start + 4 * i + 3] = (char)u;
  }
  out.resize((out.size() + 2879) / 2880 * 2880, '\0');
}
// Samples every `sampling`-th row and column (the first nNy = Ny/sampling rows
// and nNx = Nx/sampling columns).
std::vector<double> subsample(const double* data, int Nx, int Ny, int sampling, int* nNx, int* nNy) {
  *nNx = Nx / sampling;
  *nNy = Ny / sampling;
  std::vector<double> o((size_t)*nNx * *nNy);
  for (int i = 0; i < *nNy; ++i)
    for (int j = 0; j < *nNx; ++j) o[(size_t)i * *nNx + j] = data[(size_t)i * sampling * Nx + (size_t)j * sampling];
  return o;
}
}  // namespace

// Image::writeImageFITS (image.cpp:20-62): the sub-sampled image as FLOAT_IMG,
// rows flipped (first stored row = last image row), in the primary HDU and
// again in an image extension named NEW-EXTENSION.  The reference gives the
// extension the axes (nNy, nNy); that equals (nNx, nNy) for square images,
// which is what is written for all.
int mlc_write_fits(const char* path, const double* data, int Nx, int Ny, int sampling) {
  if (!path || !data || Nx <= 0 || Ny <= 0 || sampling <= 0 || Nx / sampling < 1 || Ny / sampling < 1)
    return 1;
  int nNx, nNy;
  const std::vector<double> o = subsample(data, Nx, Ny, sampling, &nNx, &nNy);
  std::vector<float> arr(o.size());
  for (int i = 0; i < nNy; ++i)
    for (int j = 0; j < nNx; ++j) arr[(size_t)(nNy - 1 - i) * nNx + j] = (float)o[(size_t)i * nNx + j];
  std::string out;
  fits_card(out, "SIMPLE", "T");
  fits_card(out, "BITPIX", "-32");
  fits_card(out, "NAXIS", "2");
  fits_card(out, "NAXIS1", std::to_string(nNx));
  fits_card(out, "NAXIS2", std::to_string(nNy));
  fits_card(out, "EXTEND", "T");
  fits_end(out);
  fits_data(out, arr);
  fits_card(out, "XTENSION", "'IMAGE   '");
  fits_card(out, "BITPIX", "-32");
  fits_card(out, "NAXIS", "2");
  fits_card(out, "NAXIS1", std::to_string(nNx));
  fits_card(out, "NAXIS2", std::to_string(nNy));
  fits_card(out, "PCOUNT", "0");
  fits_card(out, "GCOUNT", "1");
  fits_card(out, "EXTNAME", "'NEW-EXTENSION'");
  fits_end(out);
  fits_data(out, arr);
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  const size_t wr = std::fwrite(out.data(), 1, out.size(), f);
  std::fclose(f);
  return wr == out.size() ? 0 : 2;
}

namespace {
// One header card's value: the text between "= " and a '/' comment, or the
// quoted string.
std::string fits_value(const char* card) {
  std::string v(card + 10, 70);
  const size_t q = v.find('\'');
  if (q != std::string::npos && v.find_first_not_of(' ') == q) {
    const size_t e = v.find('\'', q + 1);
    return v.substr(q + 1, e == std::string::npos ? std::string::npos : e - q - 1);
  }
  const size_t sl = v.find('/');
  if (sl != std::string::npos) v.resize(sl);
  const size_t a = v.find_first_not_of(' '), b = v.find_last_not_of(' ');
  return a == std::string::npos ? std::string() : v.substr(a, b - a + 1);
}

double fits_be(const unsigned char* p, int bitpix) {
  uint64_t u = 0;
  const int nb = std::abs(bitpix) / 8;
  for (int i = 0; i < nb; ++i) u = (u << 8) | p[i];
  switch (bitpix) {
    case 8: return (double)(uint8_t)u;
    case 16: return (double)(int16_t)(uint16_t)u;
    case 32: return (double)(int32_t)(uint32_t)u;
    case 64: return (double)(int64_t)u;
    case -32: {
      const uint32_t w = (uint32_t)u;
      float f;
      std::memcpy(&f, &w, 4);
      return (double)f;
    }
    default: {
      double d;
      std::memcpy(&d, &u, 8);
      return d;
    }
  }
}
}  // namespace

// FITS image input (the optional map input; no reference reader): HDU `hdu`
// (0 = primary) of a basic FITS file, 2-D, BITPIX 8/16/32/64/-32/-64 with
// BSCALE/BZERO applied, returned with its rows flipped back, i.e. in the
// orientation that mlc_write_fits (image.cpp:20-62) was given.
int mlc_read_fits(const char* path, int hdu, double* data, int64_t cap, int* Nx, int* Ny) {
  if (!path || hdu < 0 || !Nx || !Ny || cap < 0 || (cap > 0 && !data)) return 1;
  FILE* f = std::fopen(path, "rb");
  if (!f) return 2;
  std::string raw;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) raw.append(buf, got);
  std::fclose(f);
  size_t pos = 0;
  for (int h = 0;; ++h) {
    int bitpix = 0, naxis = -1;
    int64_t ax[3] = {0, 0, 0}, pcount = 0, gcount = 1;
    double bscale = 1.0, bzero = 0.0;
    bool end = false;
    while (!end) {
      if (pos + 2880 > raw.size()) return 2;
      for (int c = 0; c < 36 && !end; ++c) {
        const char* card = raw.data() + pos + 80 * c;
        const std::string key = std::string(card, 8).substr(0, std::string(card, 8).find_last_not_of(' ') + 1);
        if (key == "END") {
          end = true;
          break;
        }
        if (card[8] != '=' || card[9] != ' ') continue;
        const std::string v = fits_value(card);
        try {
          if (key == "BITPIX") bitpix = std::stoi(v);
          else if (key == "NAXIS") naxis = std::stoi(v);
          else if (key == "NAXIS1") ax[0] = std::stoll(v);
          else if (key == "NAXIS2") ax[1] = std::stoll(v);
          else if (key == "NAXIS3") ax[2] = std::stoll(v);
          else if (key == "PCOUNT") pcount = std::stoll(v);
          else if (key == "GCOUNT") gcount = std::stoll(v);
          else if (key == "BSCALE") bscale = std::stod(v);
          else if (key == "BZERO") bzero = std::stod(v);
        } catch (...) {
          return 2;
        }
      }
      pos += 2880;
    }
    if (bitpix != 8 && bitpix != 16 && bitpix != 32 && bitpix != 64 && bitpix != -32 && bitpix != -64)
      return 2;
    if (naxis < 0 || naxis > 3) return 2;
    int64_t n = naxis == 0 ? 0 : 1;
    for (int i = 0; i < naxis; ++i) {
      if (ax[i] < 0) return 2;
      n *= ax[i];
    }
    const int bpv = std::abs(bitpix) / 8;
    const int64_t bytes = (int64_t)bpv * gcount * (pcount + n);
    if (h == hdu) {
      if (naxis != 2 || ax[0] < 1 || ax[1] < 1 || ax[0] > INT32_MAX || ax[1] > INT32_MAX) return 2;
      if (pos + (size_t)n * bpv > raw.size()) return 2;
      *Nx = (int)ax[0];
      *Ny = (int)ax[1];
      if (n > cap) return 3;
      const unsigned char* p = reinterpret_cast<const unsigned char*>(raw.data() + pos);
      for (int64_t i = 0; i < ax[1]; ++i)
        for (int64_t j = 0; j < ax[0]; ++j)
          data[(size_t)(ax[1] - 1 - i) * ax[0] + j] = bzero + bscale * fits_be(p + (size_t)(i * ax[0] + j) * bpv, bitpix);
      return 0;
    }
    pos += (size_t)((bytes + 2879) / 2880 * 2880);
    if (pos >= raw.size()) return 2;  // no such HDU
  }
}

// Image::writeImageBIN (image.cpp:4-18): the sub-sampled image as native FP64.
int mlc_write_bin(const char* path, const double* data, int Nx, int Ny, int sampling) {
  if (!path || !data || Nx <= 0 || Ny <= 0 || sampling <= 0 || Nx / sampling < 1 || Ny / sampling < 1)
    return 1;
  int nNx, nNy;
  const std::vector<double> o = subsample(data, Nx, Ny, sampling, &nNx, &nNy);
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  const size_t wr = std::fwrite(o.data(), sizeof(double), o.size(), f);
  std::fclose(f);
  return wr == o.size() ? 0 : 2;
}

// ------------------------------------------------------------------ source sizes
// BaseProfile size laws (profile.cpp:86-124), half-light radius in the units
// of r0 (parametric) or 10^14 cm (Shakura-Sunyaev disc).
double mlc_size_parametric(double r0, double l0, double nu, double lrest) {
  return r0 * std::pow(lrest / l0, nu);
}
double mlc_size_ss(double mbh, double fedd, double eta, double lrest) {
  return 0.0097 * std::pow(std::pow(lrest, 4.0) * std::pow(mbh, 2.0) * (fedd / eta), 1.0 / 3.0);
}
// Throughput-weighted sizes over a filter curve (trapezoid rule), and the
// throughput-weighted mean wavelength of the "vector" type (profile.cpp:55-84).
double mlc_size_parametric_band(double r0, double l0, double nu, const double* wl, const double* tp,
                                int n) {
  double reff = 0.0, norm = 0.0;
  for (int i = 1; i < n; ++i) {
    norm += (wl[i] - wl[i - 1]) * (tp[i] + tp[i - 1]) / 2.0;
    const double a = r0 * std::pow(wl[i - 1] / l0, nu), b = r0 * std::pow(wl[i] / l0, nu);
    reff += (wl[i] - wl[i - 1]) * (tp[i - 1] * a + tp[i] * b) / 2.0;
  }
  return reff / norm;
}
double mlc_size_ss_band(double mbh, double fedd, double eta, const double* wl, const double* tp, int n) {
  const double b2 = std::pow(mbh, 2.0), c = fedd / eta;
  double reff = 0.0, norm = 0.0;
  for (int i = 1; i < n; ++i) {
    norm += (wl[i] - wl[i - 1]) * (tp[i] + tp[i - 1]) / 2.0;
    const double a = 0.0097 * std::pow(std::pow(wl[i - 1], 4.0) * b2 * c, 1.0 / 3.0);
    const double b = 0.0097 * std::pow(std::pow(wl[i], 4.0) * b2 * c, 1.0 / 3.0);
    reff += (wl[i] - wl[i - 1]) * (tp[i - 1] * a + tp[i] * b) / 2.0;
  }
  return reff / norm;
}
double mlc_size_vector_band(const double* wl, const double* tp, int n) {
  double leff = 0.0, norm = 0.0;
  for (int i = 1; i < n; ++i) {
    norm += (wl[i] - wl[i - 1]) * (tp[i] + tp[i - 1]) / 2.0;
    leff += (wl[i] - wl[i - 1]) * (tp[i] * wl[i] + tp[i - 1] * wl[i - 1]) / 2.0;
  }
  return leff / norm;
}

}  // extern "C"
