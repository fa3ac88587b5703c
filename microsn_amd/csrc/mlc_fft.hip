// mlc_fft.hip -- C ABI of the map spectrum, profile and convolution entry
// points (kernels in mlc_fft_impl.hpp, instantiated by mlc_fft_f32.hip and
// mlc_fft_f64.hip).
#include <cmath>
#include <string>

#include "mlc_fft.hpp"

using namespace mlc;

namespace {

size_t elem(int precision) { return precision == MLC_F64 ? sizeof(double) : sizeof(float); }

ProfileParams make_params(int shape, double rhalf, double pix, double rin, double a, int Nx, int Ny) {
  MLC_REQUIRE(shape >= 0 && shape <= 5, "unknown profile shape");
  MLC_REQUIRE(std::isfinite(rhalf) && rhalf > 0, "rhalf must be positive and finite");
  MLC_REQUIRE(std::isfinite(pix) && pix > 0, "pixel size must be positive and finite");
  if (shape == MLC_SHAPE_GAUSSIAN_HOLE) MLC_REQUIRE(rin >= 0 && rin < rhalf, "need 0 <= Rin < rhalf");
  if (shape == MLC_SHAPE_WAVY) MLC_REQUIRE(a != 0.0, "wavy profile needs a != 0");
  ProfileParams P{};  // every byte defined: node tables compare their bytes
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
  if (m->precision == MLC_F64) fft_convolve_fused<double>(m, P, d_value, nullptr);
  else fft_convolve_fused<float>(m, P, d_value, nullptr);
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
    fft_profile(ctx, P, pp, dp);
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
    fft_check_dims(Nx, Ny);
    m = new mlc_map();
    m->ctx = ctx;
    m->precision = precision;
    m->Nx = Nx;
    m->Ny = Ny;
    m->W = Nx / 2 + 1;
    m->ldr = m->W + (m->W & 1);
    m->subtract_mean = subtract_mean != 0;
    size_t spec = 2 * elem(precision) * (size_t)Ny * m->W;
    hip_check(hipMalloc(&m->rows, 2 * elem(precision) * (size_t)Ny * m->ldr), "hipMalloc(row spectrum)");
    hip_check(hipMalloc(&m->fmap, spec), "hipMalloc(map spectrum)");
    hip_check(hipMalloc(&m->g0, spec), "hipMalloc(column scratch)");
    hip_check(hipMalloc(&m->g1, spec), "hipMalloc(column scratch)");
    hip_check(hipMalloc((void**)&m->mean, sizeof(double)), "hipMalloc(mean)");
    // the mean stays 0 unless subtract_mean: set once here, never by an update
    hip_check(hipMemsetAsync(m->mean, 0, sizeof(double), ctx->stream), "hipMemsetAsync(mean)");
    if (precision == MLC_F64) fft_map_transform<double>(m, d_map);
    else fft_map_transform<float>(m, d_map);
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
  (void)hipFree(m->rows);
  (void)hipFree(m->fmap);
  (void)hipFree(m->g0);
  (void)hipFree(m->g1);
  (void)hipFree(m->ntab);
  (void)hipFree(m->mean);
  const_table_free(m->jobs);
  delete m;
}

int mlc_map_update(mlc_map* m, const double* d_map) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(d_map, "mlc_map_update: null map");
    if (m->precision == MLC_F64) fft_map_transform<double>(m, d_map);
    else fft_map_transform<float>(m, d_map);
  });
}

namespace {
// The maps of a batched call: non-null, one context, one precision.
void check_maps(int nmaps, mlc_map* const* maps, const char* what) {
  MLC_REQUIRE(maps, std::string(what) + ": null map list");
  for (int i = 0; i < nmaps; ++i) {
    MLC_REQUIRE(maps[i], std::string(what) + ": null map");
    MLC_REQUIRE(maps[i]->ctx == maps[0]->ctx, std::string(what) + ": maps of different contexts");
    MLC_REQUIRE(maps[i]->precision == maps[0]->precision, std::string(what) + ": maps of different precisions");
  }
}
}  // namespace

int mlc_maps_update(int nmaps, mlc_map* const* maps, const double* const* d_maps) {
  if (nmaps < 0 || (nmaps > 0 && (!maps || !maps[0]))) return MLC_ERR_ARG;
  if (nmaps == 0) return MLC_OK;
  return guarded(maps[0]->ctx, [&] {
    check_maps(nmaps, maps, "mlc_maps_update");
    MLC_REQUIRE(d_maps, "mlc_maps_update: null source list");
    for (int i = 0; i < nmaps; ++i) MLC_REQUIRE(d_maps[i], "mlc_maps_update: null map");
    // a map listed more than once ends as the per-map calls in list order
    // leave it: updated from its last source
    std::vector<mlc_map*> ms;
    std::vector<const double*> src;
    for (int i = 0; i < nmaps; ++i) {
      bool later = false;
      for (int j = i + 1; j < nmaps; ++j) later |= maps[j] == maps[i];
      if (later) continue;
      ms.push_back(maps[i]);
      src.push_back(d_maps[i]);
    }
    if (maps[0]->precision == MLC_F64) fft_maps_transform<double>(ms, src.data());
    else fft_maps_transform<float>(ms, src.data());
  });
}

int mlc_maps_convolve_nodes(int nmaps, mlc_map* const* maps, const int* counts, const int* shape,
                            const double* rhalf, double pix, const double* rin, const double* a,
                            void* const* d_nodes) {
  if (nmaps < 0 || (nmaps > 0 && (!maps || !maps[0]))) return MLC_ERR_ARG;
  if (nmaps == 0) return MLC_OK;
  return guarded(maps[0]->ctx, [&] {
    check_maps(nmaps, maps, "mlc_maps_convolve_nodes");
    MLC_REQUIRE(counts, "mlc_maps_convolve_nodes: null counts");
    // one entry per distinct map, in order of first appearance, holding the
    // nodes of every listing of it (a map's node tables live in its own
    // storage, so each map is tabled once)
    std::vector<mlc_map*> ms;
    std::vector<std::vector<ProfileParams>> Ps;
    std::vector<std::vector<void*>> dn;
    int off = 0;
    for (int i = 0; i < nmaps; ++i) {
      MLC_REQUIRE(counts[i] >= 0, "mlc_maps_convolve_nodes: negative count");
      if (counts[i] == 0) continue;
      MLC_REQUIRE(shape && rhalf && d_nodes, "mlc_maps_convolve_nodes: null pointer");
      MLC_REQUIRE(maps[i]->Nx % 4 == 0 && maps[i]->Ny % 4 == 0,
                  "mlc_maps_convolve_nodes: node maps (4 x 4 tiles) need Nx and Ny divisible by 4");
      size_t e = 0;
      while (e < ms.size() && ms[e] != maps[i]) ++e;
      if (e == ms.size()) {
        ms.push_back(maps[i]);
        Ps.emplace_back();
        dn.emplace_back();
      }
      for (int k = off; k < off + counts[i]; ++k) {
        MLC_REQUIRE(d_nodes[k], "mlc_maps_convolve_nodes: null node map");
        Ps[e].push_back(make_params(shape[k], rhalf[k], pix, rin ? rin[k] : 0.0, a ? a[k] : 1.0, maps[i]->Nx,
                                    maps[i]->Ny));
        dn[e].push_back(d_nodes[k]);
      }
      off += counts[i];
    }
    std::vector<void*> flat;
    for (auto& v : dn) flat.insert(flat.end(), v.begin(), v.end());
    if (ms.empty()) return;
    if (maps[0]->precision == MLC_F64) fft_maps_convolve_nodes<double>(ms, Ps, flat.data());
    else fft_maps_convolve_nodes<float>(ms, Ps, flat.data());
  });
}

int mlc_map_convolve(mlc_map* m, int shape, double rhalf, double pix, double rin, double a,
                     void* d_value, void* d_node) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(d_value || d_node, "mlc_map_convolve: no output");
    if (d_node)
      MLC_REQUIRE(m->Nx % 4 == 0 && m->Ny % 4 == 0,
                  "mlc_map_convolve: node maps (4 x 4 tiles) need Nx and Ny divisible by 4");
    ProfileParams P = make_params(shape, rhalf, pix, rin, a, m->Nx, m->Ny);
    if (m->precision == MLC_F64) fft_convolve_fused<double>(m, P, d_value, d_node);
    else fft_convolve_fused<float>(m, P, d_value, d_node);
  });
}

int mlc_map_convolve_nodes(mlc_map* m, int count, const int* shape, const double* rhalf, double pix,
                           const double* rin, const double* a, void* const* d_nodes) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(count >= 0, "mlc_map_convolve_nodes: negative count");
    if (count == 0) return;
    MLC_REQUIRE(shape && rhalf && d_nodes, "mlc_map_convolve_nodes: null pointer");
    MLC_REQUIRE(m->Nx % 4 == 0 && m->Ny % 4 == 0,
                "mlc_map_convolve_nodes: node maps (4 x 4 tiles) need Nx and Ny divisible by 4");
    std::vector<ProfileParams> Ps(count);
    for (int k = 0; k < count; ++k) {
      MLC_REQUIRE(d_nodes[k], "mlc_map_convolve_nodes: null node map");
      Ps[k] = make_params(shape[k], rhalf[k], pix, rin ? rin[k] : 0.0, a ? a[k] : 1.0, m->Nx, m->Ny);
    }
    if (m->precision == MLC_F64) fft_convolve_nodes<double>(m, Ps, d_nodes);
    else fft_convolve_nodes<float>(m, Ps, d_nodes);
  });
}

int mlc_map_convolve_plain(mlc_map* m, int shape, double rhalf, double pix, double rin, double a,
                           void* d_value, void* d_node) {
  if (!m) return MLC_ERR_ARG;
  return guarded(m->ctx, [&] {
    MLC_REQUIRE(d_value || d_node, "mlc_map_convolve_plain: no output");
    if (d_node)
      MLC_REQUIRE(m->Nx % 4 == 0 && m->Ny % 4 == 0,
                  "mlc_map_convolve_plain: node maps (4 x 4 tiles) need Nx and Ny divisible by 4");
    ProfileParams P = make_params(shape, rhalf, pix, rin, a, m->Nx, m->Ny);
    if (m->precision == MLC_F64) fft_convolve_plain<double>(m, P, d_value, d_node);
    else fft_convolve_plain<float>(m, P, d_value, d_node);
  });
}

int mlc_fft_r2c(mlc_ctx* ctx, int precision, const void* d_in, void* d_out, int Nx, int Ny) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_fft_r2c: bad precision");
    MLC_REQUIRE(d_in && d_out, "mlc_fft_r2c: null pointer");
    fft_check_dims(Nx, Ny);
    if (precision == MLC_F64) fft_r2c_2d<double>(ctx, (const double*)d_in, (cpx<double>*)d_out, Nx, Ny);
    else fft_r2c_2d<float>(ctx, (const float*)d_in, (cpx<float>*)d_out, Nx, Ny);
  });
}

int mlc_fft_c2r(mlc_ctx* ctx, int precision, const void* d_in, void* d_out, int Nx, int Ny) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_fft_c2r: bad precision");
    MLC_REQUIRE(d_in && d_out, "mlc_fft_c2r: null pointer");
    fft_check_dims(Nx, Ny);
    const int W = Nx / 2 + 1;
    size_t bytes = 2 * elem(precision) * (size_t)Ny * W;
    Workspace ws(ctx);
    int it = ws.add(bytes);
    ws.commit();
    void* tmp = ws.get<void>(it);
    hip_check(hipMemcpyAsync(tmp, d_in, bytes, hipMemcpyDeviceToDevice, ctx->stream),
              "hipMemcpyAsync(D2D)");
    if (precision == MLC_F64) fft_c2r_2d<double>(ctx, (cpx<double>*)tmp, (double*)d_out, Nx, Ny);
    else fft_c2r_2d<float>(ctx, (cpx<float>*)tmp, (float*)d_out, Nx, Ny);
  });
}

}  // extern "C"
