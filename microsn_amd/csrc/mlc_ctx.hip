// mlc_ctx.hip -- context, memory helpers, per-kernel timing, and the two
// element-wise kernels of the map side: GERLUMPH normalisation and point
// sampling.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "mlc_internal.hpp"

using namespace mlc;

namespace mlc {

void timing_flush(mlc_ctx* ctx) {
  if (ctx->pending.empty()) return;
  for (auto& p : ctx->pending) {
    hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
    auto& a = ctx->acc[p.name];
    a.ms += ms;
    a.n += 1;
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
}

int fill_grid(mlc_ctx* ctx, int per_cu) {
  int cus = 0;
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
            "hipDeviceGetAttribute(CU count)");
  return std::max(1, cus) * per_cu;
}

const void* const_table(mlc_ctx* ctx, int slot, const void* host, size_t bytes) {
  return const_table_in(ctx, ctx->tables[slot], host, bytes);
}

void const_table_free(mlc_ctx::ConstTable& t) {
  if (t.dev) (void)hipFree(t.dev);
  t.dev = nullptr;
  t.cap = 0;
  t.host.clear();
}

const void* const_table_in(mlc_ctx* ctx, mlc_ctx::ConstTable& t, const void* host, size_t bytes) {
  if (t.dev && t.host.size() == bytes && std::memcmp(t.host.data(), host, bytes) == 0) return t.dev;
  if (bytes > t.cap) {
    if (t.dev) {  // kernels queued earlier may still read it
      hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(table)");
      hip_check(hipFree(t.dev), "hipFree(table)");
      t.dev = nullptr;
      t.cap = 0;
    }
    hip_check(hipMalloc(&t.dev, std::max<size_t>(bytes, 256)), "hipMalloc(table)");
    t.cap = std::max<size_t>(bytes, 256);
  }
  h2d(ctx, t.dev, host, bytes);
  ++ctx->generation;  // a graph captured before this call would read the new table
  t.host.assign(static_cast<const char*>(host), static_cast<const char*>(host) + bytes);
  return t.dev;
}

int resident_blocks(const void* kernel, int threads) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({kernel, threads});
  if (it != cache.end()) return it->second;
  int nb = 0;
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0),
            "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  nb = std::max(1, nb);
  cache[{kernel, threads}] = nb;
  return nb;
}

}  // namespace mlc

// ---------------------------------------------------------------- kernels

// magnification_map.cpp:33-38: data[i] = imap[i]*factor/muth, FP64, in that
// order of operations.
__global__ void k_gerlumph_normalise(const int32_t* __restrict__ counts, int64_t npix,
                                     double factor, double muth, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix;
       i += (int64_t)gridDim.x * blockDim.x) {
    out[i] = (double)counts[i] * factor / muth;
  }
}

// kernel_sample gpu_functions.cu:604-620: one location per thread.
// A location outside the map (which the reference would read out of bounds)
// yields NaN instead.
template <class T>
__global__ void k_sample_points(const T* __restrict__ conv, int Nx, int Ny, int k, int Nprof,
                                int Nloc, const int* __restrict__ lx, const int* __restrict__ ly,
                                double* __restrict__ LC, double* __restrict__ DLC) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= Nloc) return;
  unsigned x = (unsigned)lx[id], y = (unsigned)ly[id];
  double val = __builtin_nan("");
  if (x < (unsigned)Nx && y < (unsigned)Ny) val = (double)conv[(size_t)y * (size_t)Nx + x];
  LC[(size_t)id * Nprof + k] = val;
  if (k > 0) DLC[(size_t)id * (Nprof - 1) + k - 1] = val - LC[(size_t)id * Nprof + k - 1];
}

namespace mlc {
void make_spread_streams(mlc_ctx* ctx) {
  if (ctx->spread[0]) return;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(ctx->stream, &st), "hipStreamIsCapturing");
  if (st != hipStreamCaptureStatusNone) return;
  if (!ctx->fork_ev)
    hip_check(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming), "hipEventCreate(fork)");
  for (int i = 0; i < mlc_ctx::kSpread; ++i) {
    // default priority, like every other stream of the library
    hip_check(hipStreamCreateWithFlags(&ctx->spread[i], hipStreamNonBlocking), "hipStreamCreateWithFlags(spread)");
    hip_check(hipEventCreateWithFlags(&ctx->spread_ev[i], hipEventDisableTiming), "hipEventCreate(spread)");
  }
}

void run_spread(mlc_ctx* ctx, int n, const std::function<void(int)>& work, const std::vector<double>* weight) {
  if (n > 1) make_spread_streams(ctx);
  // with per-kernel timing on, in order: the kernels' own durations, not
  // those of kernels sharing the chip
  if (n <= 1 || !ctx->spread[0] || ctx->timing) {
    for (int i = 0; i < n; ++i) work(i);
    return;
  }
  const hipStream_t main = ctx->stream;
  const int ns = std::min(n, (int)mlc_ctx::kSpread);
  hip_check(hipEventRecord(ctx->fork_ev, main), "hipEventRecord(fork)");
  for (int s = 0; s < ns; ++s) hip_check(hipStreamWaitEvent(ctx->spread[s], ctx->fork_ev, 0), "hipStreamWaitEvent(fork)");
  auto join = [&] {
    ctx->stream = main;
    for (int s = 0; s < ns; ++s) {
      (void)hipEventRecord(ctx->spread_ev[s], ctx->spread[s]);
      (void)hipStreamWaitEvent(main, ctx->spread_ev[s], 0);
    }
  };
  // the stream of each item: round robin, or with weights the heaviest
  // items first, each to the stream with the least weight so far
  std::vector<int> order(n), on(n);
  for (int i = 0; i < n; ++i) {
    order[i] = i;
    on[i] = i % ns;
  }
  if (weight && (int)weight->size() == n) {
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return (*weight)[x] > (*weight)[y]; });
    std::vector<double> load(ns, 0.0);
    for (int i : order) {
      const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      on[i] = s;
      load[s] += (*weight)[i];
    }
  }
  try {
    for (int i : order) {
      ctx->stream = ctx->spread[on[i]];
      work(i);
    }
  } catch (...) {
    join();
    throw;
  }
  ctx->stream = main;
  for (int s = 0; s < ns; ++s) {
    hip_check(hipEventRecord(ctx->spread_ev[s], ctx->spread[s]), "hipEventRecord(spread)");
    hip_check(hipStreamWaitEvent(main, ctx->spread_ev[s], 0), "hipStreamWaitEvent(spread)");
  }
}

void run_beside(mlc_ctx* ctx, const std::function<void()>& beside, const std::function<void()>& main_work) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(ctx->stream, &st), "hipStreamIsCapturing");
  if (!ctx->side && st == hipStreamCaptureStatusNone) {
    // default priority, like every other stream of the library
    hip_check(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking), "hipStreamCreateWithFlags(side)");
    if (!ctx->fork_ev)
      hip_check(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming), "hipEventCreate(fork)");
    hip_check(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming), "hipEventCreate(join)");
  }
  if (!ctx->side) {
    beside();
    main_work();
    return;
  }
  const hipStream_t main = ctx->stream;
  hip_check(hipEventRecord(ctx->fork_ev, main), "hipEventRecord(fork)");
  hip_check(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0), "hipStreamWaitEvent(fork)");
  ctx->stream = ctx->side;
  try {
    beside();
  } catch (...) {
    ctx->stream = main;
    (void)hipEventRecord(ctx->join_ev, ctx->side);
    (void)hipStreamWaitEvent(main, ctx->join_ev, 0);
    throw;
  }
  ctx->stream = main;
  hip_check(hipEventRecord(ctx->join_ev, ctx->side), "hipEventRecord(join)");
  try {
    main_work();
  } catch (...) {
    (void)hipStreamWaitEvent(main, ctx->join_ev, 0);
    throw;
  }
  hip_check(hipStreamWaitEvent(main, ctx->join_ev, 0), "hipStreamWaitEvent(join)");
}

void sample_points(mlc_ctx* ctx, int precision, const void* conv, int Nx, int Ny, int k,
                   int Nprof, int Nloc, const int* lx, const int* ly, double* LC, double* DLC) {
  if (Nloc <= 0) return;
  dim3 grid(cdiv(Nloc, 256)), block(256);
  if (precision == MLC_F64)
    MLC_LAUNCH(ctx, "sample_points", k_sample_points<double>, grid, block, 0,
               (const double*)conv, Nx, Ny, k, Nprof, Nloc, lx, ly, LC, DLC);
  else
    MLC_LAUNCH(ctx, "sample_points", k_sample_points<float>, grid, block, 0, (const float*)conv,
               Nx, Ny, k, Nprof, Nloc, lx, ly, LC, DLC);
}
}  // namespace mlc

// ---------------------------------------------------------------- C ABI

extern "C" {

int mlc_version(void) { return 1; }

int mlc_ctx_create(int device, void* stream, mlc_ctx** out) {
  if (!out) return MLC_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return MLC_ERR_HIP;
  }
  if (device < 0 || device >= n) return MLC_ERR_ARG;
  mlc_ctx* c = new (std::nothrow) mlc_ctx();
  if (!c) return MLC_ERR_ALLOC;
  c->device = device;
  c->stream = (hipStream_t)stream;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return MLC_ERR_HIP;
  }
  *out = c;
  return MLC_OK;
}

void mlc_ctx_destroy(mlc_ctx* ctx) {
  if (!ctx) return;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto& kv : ctx->twiddles) (void)hipFree(kv.second);
  for (auto& kv : ctx->tables) (void)hipFree(kv.second.dev);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  for (int i = 0; i < mlc_ctx::kSpread; ++i) {
    if (ctx->spread[i]) {
      (void)hipStreamSynchronize(ctx->spread[i]);
      (void)hipStreamDestroy(ctx->spread[i]);
    }
    if (ctx->spread_ev[i]) (void)hipEventDestroy(ctx->spread_ev[i]);
  }
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (prev >= 0) (void)hipSetDevice(prev);
  delete ctx;
}

const char* mlc_last_error(const mlc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mlc_ctx_set_stream(mlc_ctx* ctx, void* stream) {
  return guarded(ctx, [&] { ctx->stream = (hipStream_t)stream; });
}

int mlc_ctx_generation(const mlc_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return MLC_ERR_ARG;
  *out = ctx->generation;
  return MLC_OK;
}

int mlc_ctx_reserve(mlc_ctx* ctx, size_t bytes) {
  return guarded(ctx, [&] {
    Workspace ws(ctx);
    ws.add(bytes);
    ws.commit();
  });
}

int mlc_ctx_sync(mlc_ctx* ctx) {
  return guarded(ctx, [&] { sync(ctx); });
}

int mlc_malloc(mlc_ctx* ctx, size_t bytes, void** d_ptr) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(d_ptr, "mlc_malloc: null output pointer");
    *d_ptr = nullptr;
    hip_check(hipMalloc(d_ptr, bytes ? bytes : 1), "hipMalloc");
  });
}

int mlc_free(mlc_ctx* ctx, void* d_ptr) {
  return guarded(ctx, [&] {
    if (d_ptr) {
      sync(ctx);
      hip_check(hipFree(d_ptr), "hipFree");
    }
  });
}

int mlc_memcpy_h2d(mlc_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
  return guarded(ctx, [&] {
    h2d(ctx, d_dst, src, bytes);
    sync(ctx);
  });
}

int mlc_memcpy_d2h(mlc_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
  return guarded(ctx, [&] {
    d2h(ctx, dst, d_src, bytes);
    sync(ctx);
  });
}

int mlc_memset(mlc_ctx* ctx, void* d_dst, int value, size_t bytes) {
  return guarded(ctx, [&] {
    if (bytes) hip_check(hipMemsetAsync(d_dst, value, bytes, ctx->stream), "hipMemsetAsync");
  });
}

int mlc_mem_info(mlc_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(free_bytes && total_bytes, "mlc_mem_info: null output");
    hip_check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
  });
}

int mlc_timing_enable(mlc_ctx* ctx, int on) {
  return guarded(ctx, [&] {
    if (!on) timing_flush(ctx);
    ctx->timing = on != 0;
  });
}

int mlc_timing_reset(mlc_ctx* ctx) {
  return guarded(ctx, [&] {
    timing_flush(ctx);
    ctx->acc.clear();
  });
}

int mlc_timing_get(mlc_ctx* ctx, const char* kernel, double* total_ms, long long* launches) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(kernel && total_ms && launches, "mlc_timing_get: null argument");
    timing_flush(ctx);
    auto it = ctx->acc.find(kernel);
    *total_ms = it == ctx->acc.end() ? 0.0 : it->second.ms;
    *launches = it == ctx->acc.end() ? 0 : it->second.n;
  });
}

int mlc_timing_names(mlc_ctx* ctx, char* buf, size_t cap) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(buf && cap > 0, "mlc_timing_names: null buffer");
    timing_flush(ctx);
    std::string s;
    for (auto& kv : ctx->acc) s += kv.first + "\n";
    size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  });
}

int mlc_gerlumph_normalise(mlc_ctx* ctx, const int32_t* d_counts, int64_t npix, double avgmu,
                           double avgN, double k, double g, double* d_map) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(npix >= 0, "mlc_gerlumph_normalise: negative size");
    MLC_REQUIRE(npix == 0 || (d_counts && d_map), "mlc_gerlumph_normalise: null pointer");
    MLC_REQUIRE(avgN != 0.0, "mlc_gerlumph_normalise: avgN == 0");
    double factor = std::fabs(avgmu / avgN);
    double den = std::pow(1.0 - k, 2) - std::pow(g, 2);
    MLC_REQUIRE(den != 0.0, "mlc_gerlumph_normalise: (1-k)^2 - g^2 == 0");
    double muth = std::fabs(1.0 / den);
    if (npix == 0) return;
    int grid = std::min<int64_t>(cdiv(npix, 256), fill_grid(ctx, 8));
    MLC_LAUNCH(ctx, "gerlumph_normalise", k_gerlumph_normalise, dim3(grid), dim3(256), 0, d_counts,
               npix, factor, muth, d_map);
  });
}

int mlc_sample_points(mlc_ctx* ctx, int precision, const void* d_conv, int Nx, int Ny, int k,
                      int Nprof, int Nloc, const int* d_loc_x, const int* d_loc_y, double* d_LC,
                      double* d_DLC) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(precision == MLC_F32 || precision == MLC_F64, "mlc_sample_points: bad precision");
    MLC_REQUIRE(Nx > 0 && Ny > 0 && Nprof > 0 && k >= 0 && k < Nprof && Nloc >= 0,
                "mlc_sample_points: bad sizes");
    MLC_REQUIRE(Nloc == 0 || (d_conv && d_loc_x && d_loc_y && d_LC), "mlc_sample_points: null pointer");
    MLC_REQUIRE(k == 0 || d_DLC, "mlc_sample_points: DLC required for k > 0");
    sample_points(ctx, precision, d_conv, Nx, Ny, k, Nprof, Nloc, d_loc_x, d_loc_y, d_LC, d_DLC);
  });
}

}  // extern "C"
