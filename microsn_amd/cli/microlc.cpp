// microlc -- the reference's command-line driver (simulating_lcurves/src/
// driver.cpp:19-270) on top of libmlc.so: one JSON input file, the same
// stdout banners, the same output files in the working directory
// (mpd_{A,B,R}.dat, binned_exp.dat, binned_N.dat, lcurves{A,B}.dat) and the
// same last line "GPU: <integral>".
//
// Differences from the reference, all on error or output paths:
//   * input errors exit with status 1 (driver.cpp:92 returns 0);
//   * lcurves{A,B}.dat hold the computed light curves (the reference writes
//     an unfilled host buffer unless its CPU comparison is switched on);
//   * the maps are read from the directories given by "maps": {"A": dir,
//     "B": dir} when present, else from $MLC_MAP_PATH/<id> with the
//     reference's ids 12123 and 11156 (driver.cpp:133-134);
//   * optional keys "M" (two masses), "V" (velocity) and "precision"
//     ("f64", the default, or "f32") replace constants of driver.cpp:37,195-196.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mlc.h"
#include "../../include/mlc_host.h"
#include "json_min.hpp"

namespace {

const char* kShapes[] = {"uniform", "gaussian", "gaussian_hole", "thermal_hole", "wavy", "exponential"};

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

mlc_ctx* g_ctx = nullptr;

void check(int rc, const char* what) {
  if (rc != MLC_OK) throw Fail(std::string(what) + ": " + (g_ctx ? mlc_last_error(g_ctx) : "error"));
}

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  check(mlc_malloc(g_ctx, n * sizeof(T), &p), "mlc_malloc");
  return (T*)p;
}

int shape_id(const std::string& s) {
  for (int i = 0; i < 6; ++i)
    if (s == kShapes[i]) return i;
  throw Fail("unknown profile shape \"" + s + "\"");
}

void write_pairs(const std::string& name, const std::vector<double>& a, const std::vector<double>& b) {
  FILE* fh = std::fopen(name.c_str(), "w");
  if (!fh) throw Fail("cannot write " + name);
  for (size_t i = 0; i < a.size(); ++i) std::fprintf(fh, "%11.6e %11.6e\n", a[i], b[i]);
  std::fclose(fh);
}

void stars() { std::cout << std::string(100, '*') << std::endl; }

struct Map {
  int Nx = 0;
  double width = 0.0;
  double* d = nullptr;  // device FP64 magnifications
};

// MagnificationMap(id, Rein = 1) (magnification_map.cpp:7-40).
Map load_map(const std::string& dir) {
  double avgmu, avgN, width, k, g, s;
  int Nx;
  int rc = mlc_gerlumph_read(dir.c_str(), &avgmu, &avgN, &Nx, &width, &k, &g, &s, nullptr, 0);
  if (rc) throw Fail("cannot read map metadata in " + dir);
  std::vector<int32_t> counts((size_t)Nx * Nx);
  rc = mlc_gerlumph_read(dir.c_str(), &avgmu, &avgN, &Nx, &width, &k, &g, &s, counts.data(),
                         (int64_t)counts.size());
  if (rc) throw Fail("cannot read map.bin in " + dir);
  int32_t* dc = dalloc<int32_t>(counts.size());
  check(mlc_memcpy_h2d(g_ctx, dc, counts.data(), counts.size() * sizeof(int32_t)), "upload map");
  Map m;
  m.Nx = Nx;
  m.width = width;
  m.d = dalloc<double>(counts.size());
  check(mlc_gerlumph_normalise(g_ctx, dc, (int64_t)counts.size(), avgmu, avgN, k, g, m.d),
        "mlc_gerlumph_normalise");
  check(mlc_free(g_ctx, dc), "mlc_free");
  return m;
}

int run(const std::string& input_path) {
  std::ifstream fin(input_path);
  if (!fin) throw Fail("cannot open " + input_path);
  std::stringstream buf;
  buf << fin.rdbuf();
  const mlcjson::Value in = mlcjson::parse(buf.str());

  // ---- inputs and checks (driver.cpp:41-120), all before any device work
  const int Nprof = (int)in["sizes"].size();
  if (Nprof < 1) throw Fail("\"sizes\" is empty");
  std::vector<double> sizes(Nprof);
  for (int i = 0; i < Nprof; ++i) sizes[i] = in["sizes"][i].asDouble();
  const std::string shape = in["shape"].asString();
  const int sid = shape_id(shape);
  std::printf("Number of profiles set to: %d (%s)\n", Nprof, shape.c_str());

  const mlcjson::Value& loc = in["locations"];
  const std::string ltype = loc["type"].asString();
  const int Nmap = loc["Nmap"].asInt();
  const double ss = loc["ss"].asDouble();
  const int Noff = (int)std::ceil((Nmap / ss) * sizes.back());
  int Nloc;
  std::vector<int> x, y;
  if (ltype == "grid") {
    const int Nside = loc["Nside"].asInt();
    Nloc = Nside * Nside;
    x.resize(Nloc);
    y.resize(Nloc);
    if (mlc_locations_on_grid(Nside, Nmap, Noff, x.data(), y.data())) throw Fail("bad grid locations");
  } else {
    Nloc = loc["Nloc"].asInt();
    x.resize(Nloc);
    y.resize(Nloc);
    if (mlc_locations_random(Nloc, Nmap, Noff, loc["seed"].asInt(), x.data(), y.data()))
      throw Fail("bad random locations");
  }
  std::printf("Sampling from %d locations per map (%s).\n", Nloc, ltype.c_str());
  std::printf("Will loop %.2f times (round up) to cover all location pairs.\n", Nloc / 1024.0);

  const mlcjson::Value& data = in["data"][0];
  const int Nd = (int)data["d"].size();
  if (Nd >= 50) throw Fail("Number of data points is too large to fit in register memory!");
  std::printf("Data length: %d\n", Nd);
  std::vector<double> d(Nd), s(Nd), t(Nd);
  for (int i = 0; i < Nd; ++i) {
    d[i] = data["d"][i].asDouble();
    s[i] = data["s"][i].asDouble();
    t[i] = data["t"][i].asDouble();
  }
  const double Dfac = std::sqrt(in["DS"].asDouble() * in["DLS"].asDouble() / in["DL"].asDouble());
  const mlcjson::Value& mp = in["mpds"];
  const int Nbins_mpd = mp["mpd"]["Nbins"].asInt(), Nbins_ratio = mp["ratio"]["Nbins"].asInt();
  const double min_mpd = mp["mpd"]["min"].asDouble(), max_mpd = mp["mpd"]["max"].asDouble();
  const double min_ratio = mp["ratio"]["min"].asDouble(), max_ratio = mp["ratio"]["max"].asDouble();
  double MA = 0.4, MB = 0.5, V = 1.0;
  if (in.has("M")) {
    MA = in["M"][0].asDouble();
    MB = in["M"][1].asDouble();
  }
  if (in.has("V")) V = in["V"][0].asDouble();
  int precision = MLC_F64;
  if (in.has("precision")) {
    const std::string pr = in["precision"].asString();
    if (pr == "f32") precision = MLC_F32;
    else if (pr != "f64") throw Fail("precision must be \"f64\" or \"f32\"");
  }
  std::string dirA, dirB;
  if (in.has("maps")) {
    dirA = in["maps"]["A"].asString();
    dirB = in["maps"]["B"].asString();
  } else {
    const char* mp_env = std::getenv("MLC_MAP_PATH");
    const std::string base = mp_env ? std::string(mp_env) : std::string("./");
    dirA = base + "12123";
    dirB = base + "11156";
  }

  // ---- device
  check(mlc_ctx_create(0, nullptr, &g_ctx), "mlc_ctx_create");
  size_t free_b = 0, total_b = 0;
  check(mlc_mem_info(g_ctx, &free_b, &total_b), "mlc_mem_info");
  // driver.cpp:110-111 prints get_total_gpu_mem(), which is in MiB
  std::printf("Total GPU memory: %ld (bytes)\n", (long)(total_b / (1024 * 1024)));
  Map A = load_map(dirA), B = load_map(dirB);
  if (A.Nx != Nmap || B.Nx != Nmap)
    std::fprintf(stderr, "warning: the maps have %d and %d pixels a side, \"Nmap\" says %d\n", A.Nx, B.Nx,
                 Nmap);
  int *dx = dalloc<int>(Nloc), *dy = dalloc<int>(Nloc);
  check(mlc_memcpy_h2d(g_ctx, dx, x.data(), sizeof(int) * Nloc), "upload x");
  check(mlc_memcpy_h2d(g_ctx, dy, y.data(), sizeof(int) * Nloc), "upload y");
  const size_t nlc = (size_t)Nloc * Nprof, ndlc = (size_t)Nloc * (Nprof > 1 ? Nprof - 1 : 1);
  double *LCA = dalloc<double>(nlc), *DLCA = dalloc<double>(ndlc);
  double *LCB = dalloc<double>(nlc), *DLCB = dalloc<double>(ndlc);

  std::cout << "CONVOLUTIONS" << std::endl;
  stars();
  check(mlc_expanding_source(g_ctx, precision, A.d, A.Nx, A.Nx, A.width, sizes.data(), Nprof, sid, Nloc, dx,
                             dy, LCA, DLCA),
        "expanding_source (map A)");
  for (int k = 1; k < Nprof; ++k) std::cout << "Profile: " << k << std::endl;
  check(mlc_expanding_source(g_ctx, precision, B.d, B.Nx, B.Nx, B.width, sizes.data(), Nprof, sid, Nloc, dx,
                             dy, LCB, DLCB),
        "expanding_source (map B)");
  for (int k = 1; k < Nprof; ++k) std::cout << "Profile: " << k << std::endl;

  std::cout << "MPDS" << std::endl;
  stars();
  std::vector<double> bA(Nbins_mpd), cA(Nbins_mpd), bB(Nbins_mpd), cB(Nbins_mpd);
  std::vector<double> bR(Nbins_ratio), cR(Nbins_ratio);
  if (mlc_mpd_bins(Nbins_mpd, min_mpd, max_mpd, 1, bA.data())) throw Fail("bad \"mpd\" bins");
  bB = bA;
  check(mlc_binned_mpd(g_ctx, A.d, (int64_t)A.Nx * A.Nx, Nbins_mpd, min_mpd, max_mpd, bA.data(), cA.data()),
        "getBinnedMpd (map A)");
  check(mlc_binned_mpd(g_ctx, B.d, (int64_t)B.Nx * B.Nx, Nbins_mpd, min_mpd, max_mpd, bB.data(), cB.data()),
        "getBinnedMpd (map B)");
  mlc_mpd_normalize(Nbins_mpd, bA.data(), cA.data());
  mlc_mpd_normalize(Nbins_mpd, bB.data(), cB.data());
  if (mlc_mpd_bins(Nbins_ratio, min_ratio, max_ratio, 1, bR.data())) throw Fail("bad \"ratio\" bins");
  if (mlc_mpd_divide_by(Nbins_mpd, bA.data(), cA.data(), Nbins_mpd, bB.data(), cB.data(), Nbins_ratio,
                        bR.data(), cR.data()))
    throw Fail("MPD ratio failed");
  write_pairs("mpd_A.dat", bA, cA);
  write_pairs("mpd_B.dat", bB, cB);
  write_pairs("mpd_R.dat", bR, cR);

  const size_t npair = (size_t)Nloc * Nloc;
  uint32_t* sorted_ind = dalloc<uint32_t>(npair);
  std::vector<uint32_t> lower_ind(Nbins_ratio), n_per_bin(Nbins_ratio);
  check(mlc_setup_integral(g_ctx, LCA, LCB, Nloc, Nprof, bR.data(), Nbins_ratio, sorted_ind,
                           lower_ind.data(), n_per_bin.data()),
        "setup_integral");

  std::cout << "CHI-SQUARED" << std::endl;
  stars();
  std::vector<int> indA(Nd), indB(Nd);
  std::vector<double> facA(Nd), facB(Nd), nd(Nd), ns(Nd);
  int Njp = 0;
  double const_term = 0.0;
  if (mlc_setup_chi2(MA, MB, V, Dfac, Nd, d.data(), t.data(), s.data(), Nprof, sizes.data(), 1, indA.data(),
                     indB.data(), facA.data(), facB.data(), nd.data(), ns.data(), &Njp, &const_term))
    throw Fail("setup_chi2_calculation failed");
  double* chi2 = dalloc<double>(npair);
  check(mlc_calculate_chi2(g_ctx, Njp, indA.data(), indB.data(), facA.data(), facB.data(), nd.data(), ns.data(),
                           sorted_ind, LCA, DLCA, LCB, DLCB, Nloc, Nprof, chi2),
        "calculate_chi2");
  std::vector<double> binned_chi2(Nbins_ratio), binned_exp(Nbins_ratio);
  check(mlc_bin_chi2(g_ctx, chi2, (int64_t)npair, lower_ind.data(), n_per_bin.data(), Nbins_ratio,
                     binned_chi2.data(), binned_exp.data()),
        "bin_chi2");
  const double integral = mlc_trapezium(Nbins_ratio, binned_exp.data(), cR.data(), bR.data());
  std::cout << "GPU: " << integral << std::endl;

  FILE* fh = std::fopen("binned_exp.dat", "w");
  FILE* fh2 = std::fopen("binned_N.dat", "w");
  if (!fh || !fh2) throw Fail("cannot write binned_*.dat");
  for (int i = 0; i < Nbins_ratio; ++i) {
    std::fprintf(fh, "%11.6e %11.6e\n", bR[i], binned_exp[i]);
    std::fprintf(fh2, "%11.6e %d\n", bR[i], (int)n_per_bin[i]);
  }
  std::fclose(fh);
  std::fclose(fh2);

  std::vector<double> hA(nlc), hB(nlc);
  check(mlc_memcpy_d2h(g_ctx, hA.data(), LCA, nlc * sizeof(double)), "download LCA");
  check(mlc_memcpy_d2h(g_ctx, hB.data(), LCB, nlc * sizeof(double)), "download LCB");
  fh = std::fopen("lcurvesA.dat", "w");
  fh2 = std::fopen("lcurvesB.dat", "w");
  if (!fh || !fh2) throw Fail("cannot write lcurves*.dat");
  for (int i = 0; i < Nloc; ++i) {
    for (int j = 0; j < Nprof; ++j) {
      std::fprintf(fh, " %f", hA[(size_t)i * Nprof + j]);
      std::fprintf(fh2, " %f", hB[(size_t)i * Nprof + j]);
    }
    std::fprintf(fh, "\n");
    std::fprintf(fh2, "\n");
  }
  std::fclose(fh);
  std::fclose(fh2);

  for (void* p : {(void*)A.d, (void*)B.d, (void*)dx, (void*)dy, (void*)LCA, (void*)DLCA, (void*)LCB,
                  (void*)DLCB, (void*)sorted_ind, (void*)chi2})
    mlc_free(g_ctx, p);
  mlc_ctx_destroy(g_ctx);
  g_ctx = nullptr;
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc != 2) {
    std::cout << "Error: You need to provide the name of a .json file that contains all the input variables!"
              << std::endl;
    return 1;
  }
  try {
    return run(argv[1]);
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    if (g_ctx) mlc_ctx_destroy(g_ctx);
    return 1;
  }
}
