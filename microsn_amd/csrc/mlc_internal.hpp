// mlc_internal.hpp -- shared host-side plumbing of libmlc.so: the context
// (stream, last error, workspace, per-kernel event timing), status handling
// and the launch wrapper used by every kernel file.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mlc.h"

namespace mlc {

// Error carried from the depths of a call to the C-ABI boundary, where it is
// turned into a status code and the context's last-error text.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    // clear the sticky per-thread error so the next call starts clean
    (void)hipGetLastError();
    fail(e == hipErrorOutOfMemory ? MLC_ERR_ALLOC : MLC_ERR_HIP,
         std::string(what) + ": " + hipGetErrorString(e));
  }
}

#define MLC_REQUIRE(cond, msg)                       \
  do {                                               \
    if (!(cond)) ::mlc::fail(MLC_ERR_ARG, (msg));    \
  } while (0)

}  // namespace mlc

struct mlc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // growable device workspace, carved into aligned slices per call
  void* ws = nullptr;
  size_t ws_bytes = 0;

  // bumped whenever device memory that a captured graph may point at is
  // reallocated (workspace, a map's node tables) or a device table is
  // rewritten in place (const_table); a graph captured at generation g is
  // valid while the generation is still g
  uint64_t generation = 0;

  // launch-shape tuner of the light-curve band kernels (mlc_tune_*): mode,
  // the forced shape, and the kept shape (lanes, grid multiple) per key
  int tune_mode = MLC_TUNE_AUTO;
  int tune_force_lanes = 16, tune_force_gmul = 1, tune_force_flags = 0;
  struct TunedShape {
    int lanes, gmul, per_band;
  };
  std::map<uint64_t, TunedShape> tune_cache;

  // per-kernel timing
  bool timing = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  struct Acc {
    double ms = 0.0;
    long long n = 0;
  };
  std::map<std::string, Acc> acc;

  // twiddle tables: key = (n, precision) -> device array of n complex values
  std::map<std::pair<int, int>, void*> twiddles;

  // small host tables kept on the device between calls (mlc::const_table)
  struct ConstTable {
    std::vector<char> host;
    void* dev = nullptr;
    size_t cap = 0;
  };
  std::map<int, ConstTable> tables;

  // a second stream (default priority) for work that runs beside the main
  // stream inside one call, forked and joined by the two events (created on
  // first use, outside graph capture)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // more streams (default priority) for independent per-map work of one call
  // (run_spread), with a join event each, created like `side`
  static constexpr int kSpread = 4;
  hipStream_t spread[kSpread] = {};
  hipEvent_t spread_ev[kSpread] = {};
};

namespace mlc {

// Runs `beside` on the context's side stream while `main_work` runs on the
// main stream: the side stream waits for the main stream's work so far, and
// the main stream waits for `beside` before anything issued after this call.
// Under graph capture the two become parallel branches of the graph.  The
// side stream is created on the first call that is not being captured; until
// then (and without it) the two run one after the other.  Kernels launched by
// `beside` go to the side stream through ctx->stream.
void run_beside(mlc_ctx* ctx, const std::function<void()>& beside, const std::function<void()>& main_work);

// Runs work(i) for i = 0 .. n-1, item i on spread stream i mod kSpread (with
// `weight`, heaviest first, each on the stream with the least weight so
// far): the
// spread streams wait for the main stream's work so far, and the main stream
// waits for all of them before anything issued after this call.  Items run
// concurrently, so they must not share scratch memory (per-map work on
// per-map buffers).  The streams are created on the first call that is not
// being captured; until then, for n <= 1 and while per-kernel timing is on,
// the items run in order on the main stream.  Under graph capture they
// become parallel branches.
void run_spread(mlc_ctx* ctx, int n, const std::function<void(int)>& work,
                const std::vector<double>* weight = nullptr);

// Creates the spread streams and their events if they do not exist yet and
// the context's stream is not being captured (run_spread does so itself;
// this is for a caller that spreads its work only under capture).
void make_spread_streams(mlc_ctx* ctx);

// Makes `ctx`'s device current for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    hip_check(hipGetDevice(&prev), "hipGetDevice");
    if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
  }
  ~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
  }
};

// Slices of a growable device buffer: the context workspace by default, or
// a buffer owned by another object (so that two live Workspaces never share
// memory).  Grows (free + malloc) when a call needs more than the current
// size; the previous contents are not kept.
struct Workspace {
  mlc_ctx* ctx;
  void** buf;
  size_t* cap;
  size_t used = 0;
  std::vector<std::pair<size_t, size_t>> slices;  // offset, bytes
  // fixed mode: a caller-provided buffer that is never reallocated
  bool fixed = false;
  void* fixed_buf = nullptr;
  size_t fixed_cap = 0;
  const char* fixed_what = nullptr;
  explicit Workspace(mlc_ctx* c) : ctx(c), buf(&c->ws), cap(&c->ws_bytes) {}
  Workspace(mlc_ctx* c, void** b, size_t* n) : ctx(c), buf(b), cap(n) {}
  // A buffer of `bytes` at device address `base` (256-byte aligned): commit()
  // fails with MLC_ERR_ARG when the slices do not fit, and never allocates.
  Workspace(mlc_ctx* c, void* base, size_t bytes, const char* what)
      : ctx(c), buf(&fixed_buf), cap(&fixed_cap), fixed(true), fixed_buf(base), fixed_cap(bytes),
        fixed_what(what) {}
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  // Declare a slice; returns its index.  Call commit() before get().
  int add(size_t bytes) {
    size_t off = (used + 255) & ~(size_t)255;
    slices.push_back({off, bytes});
    used = off + bytes;
    return (int)slices.size() - 1;
  }
  void commit() {
    if (fixed) {
      if (used > fixed_cap)
        fail(MLC_ERR_ARG, std::string(fixed_what) + ": this call needs " + std::to_string(used) +
                              " bytes of scratch, the static scratch holds " + std::to_string(fixed_cap));
      return;
    }
    if (used > *cap) {
      if (*buf) {
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(workspace)");
        hip_check(hipFree(*buf), "hipFree(workspace)");
        *buf = nullptr;
        *cap = 0;
      }
      size_t sz = used + (used >> 3) + 4096;
      hip_check(hipMalloc(buf, sz), "hipMalloc(workspace)");
      *cap = sz;
      ++ctx->generation;
    }
  }
  template <class T>
  T* get(int i) const {
    return reinterpret_cast<T*>(static_cast<char*>(*buf) + slices[i].first);
  }
};

// Brackets one kernel launch with HIP events when timing is on.
struct KernelTimer {
  mlc_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(mlc_ctx* c, const char* n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    a = take();
    b = take();
    hip_check(hipEventRecord(a, ctx->stream), "hipEventRecord");
  }
  ~KernelTimer() noexcept(false) {
    if (!ctx->timing || !a) return;
    hip_check(hipEventRecord(b, ctx->stream), "hipEventRecord");
    ctx->pending.push_back({name, a, b});
  }
  hipEvent_t take() {
    hipEvent_t e;
    if (!ctx->event_pool.empty()) {
      e = ctx->event_pool.back();
      ctx->event_pool.pop_back();
    } else {
      hip_check(hipEventCreate(&e), "hipEventCreate");
    }
    return e;
  }
};

void timing_flush(mlc_ctx* ctx);  // resolve pending events into ctx->acc

// Launch wrapper: timing + immediate launch-error check.
#define MLC_LAUNCH(ctx, name, kern, grid, block, shmem, ...)                     \
  do {                                                                           \
    ::mlc::KernelTimer _mlc_t((ctx), (name));                                    \
    hipLaunchKernelGGL(kern, (grid), (block), (shmem), (ctx)->stream, __VA_ARGS__); \
    ::mlc::hip_check(hipGetLastError(), name);                                   \
  } while (0)

inline void sync(mlc_ctx* ctx) { hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"); }

inline void h2d(mlc_ctx* ctx, void* d, const void* h, size_t bytes) {
  if (bytes) hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync(H2D)");
}
inline void d2h(mlc_ctx* ctx, void* h, const void* d, size_t bytes) {
  if (bytes) hip_check(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync(D2H)");
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Number of workgroups that fill the device a few times over (grid-stride kernels).
int fill_grid(mlc_ctx* ctx, int per_cu);

// A device copy of a small host table (bin edges, node jobs), one per slot:
// uploaded when its bytes differ from the previous call's, reused otherwise,
// so that a repeated call issues no host-to-device copy (and can be captured
// into a graph once it has run).
// kTableNodeJobs and kTableDcJobs: the batched map calls (mlc_maps_*)
enum ConstSlot { kTableBins = 0, kTableNodeJobs = 1, kTableBatchJobs = 2, kTableBatchMaps = 3, kTableDcJobs = 4 };
const void* const_table(mlc_ctx* ctx, int slot, const void* host, size_t bytes);
// The same for a table owned by another object (e.g. a map's node jobs, so
// that several maps used in one step each keep theirs on the device).
const void* const_table_in(mlc_ctx* ctx, mlc_ctx::ConstTable& t, const void* host, size_t bytes);
void const_table_free(mlc_ctx::ConstTable& t);

// Workgroups of `threads` lanes of `kernel` resident at once on one CU (the
// occupancy limit of its registers and LDS), at least 1; cached per kernel.
int resident_blocks(const void* kernel, int threads);

// Turns an exception into a status code and records the message.
template <class F>
int guarded(mlc_ctx* ctx, F&& f) {
  try {
    if (!ctx) return MLC_ERR_ARG;
    DeviceGuard g(ctx->device);
    f();
    return MLC_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return MLC_ERR_ALLOC;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return MLC_ERR_HIP;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return MLC_ERR_HIP;
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

}  // namespace mlc
