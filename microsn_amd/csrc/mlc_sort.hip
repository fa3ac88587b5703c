// mlc_sort.hip -- stable LSD radix sort of FP64 keys with a 32-bit index
// payload, and the bin table over the sorted keys.  Replaces
// thrust::stable_sort_by_key + thrust::upper_bound + adjacent_difference
// (gpu_functions.cu:368-397) with the ordering of std::stable_sort on '<'
// (cpu_functions.cpp:179-181).
//
// Keys are mapped to uint64 so that unsigned order equals '<' order:
// -0.0 becomes +0.0 (they compare equal), every NaN becomes the largest
// value (NaN sorts after +inf; among NaNs the original index order is kept),
// negative values are bit-inverted and positive ones get the sign bit set.
// Eight 8-bit digit passes.  The transform kernel histograms all eight
// digits; a pass whose digit is the same for every key is trivial and its
// kernels reduce to a copy.  The decision is taken on the device by every
// kernel of the pass, so the sort never waits for the host.
//
// Up to 2^21 keys: segments of 2048 keys are sorted in LDS (all eight passes
// without leaving the workgroup), then merged pairwise (stable merge path,
// ceil(log2(n/2048)) rounds).  Larger sorts run eight passes over the whole
// array.
//
// One pass over tiles of 256*R keys (256 threads x R rounds, round r holds
// keys r*256 + t, i.e. tile order = (round, thread); R = 16):
//   k_tile_hist  per-tile digit counts, digit-major [256][ntiles]
//   k_scan_tiles exclusive offsets: digit base + earlier tiles of that digit
//   k_scatter    stable rank inside the tile (wavefront match on the 8 digit
//                bits + per-wave counts in LDS), keys staged in LDS in digit
//                order, then written out in runs that are contiguous in memory.
#include <algorithm>
#include <vector>

#include "mlc_internal.hpp"

using namespace mlc;

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int64_t kSmallSort = 1 << 21;  // up to this many keys: segment sorts and merges
constexpr int kScatterRounds = 16;       // rounds of a radix pass's tile (256 keys each)
constexpr int64_t kRankWindow = (int64_t)32 << 20;  // ranks per scatter window (128 MB)

// True when every key of the pass has the same digit (block-uniform).
__device__ __forceinline__ bool pass_trivial(const unsigned* __restrict__ hist_pass, int64_t n) {
  return __syncthreads_or(hist_pass[threadIdx.x] == (unsigned)n) != 0;
}

__device__ __forceinline__ uint64_t key_of(double x) {
  uint64_t b = (uint64_t)__double_as_longlong(x);
  if (x != x) return ~0ull;                      // NaN
  if (b == 0x8000000000000000ull) b = 0;         // -0.0 -> +0.0
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ unsigned digit_of(uint64_t k, int pass) {
  return (unsigned)(k >> (8 * pass)) & 255u;
}

__global__ void k_zero_u32(unsigned* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// h[d] += 1 for the digit d of every active lane.  The digits of the first
// two distinct values among the wave's lanes are added with one atomic each
// (the high digits of keys from a narrow range take one or two values, and
// 64 lanes on one LDS address serialise), the rest one by one.
__device__ __forceinline__ void wave_count(unsigned* h, unsigned d) {
  const int lane = threadIdx.x & 63;
  bool done = false;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const uint64_t act = __ballot(!done);
    if (act == 0) break;  // uniform
    const int fl = __builtin_ctzll(act);
    const unsigned dv = __shfl(d, fl, 64);
    const uint64_t m = __ballot(!done && d == dv);
    if (lane == fl) atomicAdd(&h[dv], (unsigned)__popcll(m));
    done |= d == dv;
  }
  if (!done) atomicAdd(&h[d], 1u);
}

// Transform keys and histogram all eight digits.  hist: [8][256] uint32.
constexpr int kHistThreads = 1024;
__global__ __launch_bounds__(kHistThreads) void k_transform_hist(const double* __restrict__ in,
                                                                 uint64_t* __restrict__ out, int64_t n,
                                                                 unsigned* __restrict__ hist) {
  __shared__ unsigned h[8][256];
  for (int i = threadIdx.x; i < 8 * 256; i += kHistThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kHistThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kHistThreads) {
    uint64_t k = key_of(in[i]);
    out[i] = k;
    // the low six digits are close to uniform: plain atomics; the sign and
    // exponent digits often take one or two values
#pragma unroll
    for (int p = 0; p < 6; ++p) atomicAdd(&h[p][digit_of(k, p)], 1u);
#pragma unroll
    for (int p = 6; p < 8; ++p) wave_count(h[p], digit_of(k, p));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 8 * 256; i += kHistThreads) {
    unsigned v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// Per-tile digit counts, digit-major.  Persistent, with the tiles split over
// the XCDs as k_scatter splits them (grid a multiple of 8): the counts of
// neighbouring tiles, which share cache lines of a digit's row, are
// completed in one L2.  The next tile's keys are loaded while the current
// tile's are counted.
template <int R>
__global__ __launch_bounds__(kThreads) void k_tile_hist(const uint64_t* __restrict__ keys, int64_t n,
                                                        int pass, int ntiles,
                                                        unsigned* __restrict__ counts,
                                                        const unsigned* __restrict__ hist_pass) {
  if (pass_trivial(hist_pass, n)) return;
  __shared__ unsigned h[256];
  const int per = (ntiles + 7) >> 3, xcd = blockIdx.x & 7, nwx = gridDim.x >> 3;
  const int tend = min(ntiles, (xcd + 1) * per);
  int tile = xcd * per + (int)(blockIdx.x >> 3);
  uint64_t kr[R];  // raw keys of the next tile: used only once they are needed
  auto load = [&](int tl) {
    const int64_t base = (int64_t)tl * (kThreads * R);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = base + r * kThreads + threadIdx.x;
      kr[r] = i < n ? keys[i] : 0ull;
    }
  };
  if (tile < tend) load(tile);
  for (; tile < tend; tile += nwx) {  // uniform per workgroup
    const int64_t base = (int64_t)tile * (kThreads * R);
    unsigned d[R];
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = base + r * kThreads + threadIdx.x < n ? digit_of(kr[r], pass) : 256u;
    if (tile + nwx < tend) load(tile + nwx);
    h[threadIdx.x] = 0;
    __syncthreads();
    if (pass >= 6) {  // uniform: the sign and exponent digits
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (d[r] < 256u) wave_count(h, d[r]);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (d[r] < 256u) atomicAdd(&h[d[r]], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * ntiles + tile] = h[threadIdx.x];
    __syncthreads();  // h is cleared for the next tile
  }
}

// One workgroup per digit d: exclusive scan of the digit's tile counts
// (each thread sums a contiguous run of tiles, the 256 run totals are
// scanned in LDS) plus the digit's base, the number of keys with a smaller
// digit (a block reduction over the pass histogram).  With `staged`, the
// digit's row of counts is first copied into LDS (dynamic shared memory of
// ntiles words) with coalesced loads, scanned there and copied back, instead
// of each thread walking its run of the row in global memory.
constexpr int kScanStageTiles = 32768;  // rows up to this long are staged (128 KB)
__global__ __launch_bounds__(kThreads) void k_scan_tiles(unsigned* __restrict__ counts, int ntiles,
                                                         const unsigned* __restrict__ hist_pass,
                                                         int64_t n, int staged) {
  if (pass_trivial(hist_pass, n)) return;
  extern __shared__ unsigned rowl[];
  __shared__ unsigned s[kThreads];
  const int d = blockIdx.x, t = threadIdx.x;
  unsigned* row = counts + (size_t)d * ntiles;
  unsigned* src = row;
  if (staged) {
    constexpr int kBatch = 8;
    for (int i0 = 0; i0 < ntiles; i0 += kBatch * kThreads) {
      unsigned v[kBatch];
#pragma unroll
      for (int k = 0; k < kBatch; ++k) {
        const int i = i0 + k * kThreads + t;
        v[k] = i < ntiles ? row[i] : 0u;
      }
#pragma unroll
      for (int k = 0; k < kBatch; ++k) {
        const int i = i0 + k * kThreads + t;
        if (i < ntiles) rowl[i] = v[k];
      }
    }
    src = rowl;
  }
  // digit base
  s[t] = t < d ? hist_pass[t] : 0u;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (t < w) s[t] += s[t + w];
    __syncthreads();
  }
  const unsigned base = s[0];
  __syncthreads();
  const int seg = (ntiles + kThreads - 1) / kThreads;
  const int a = min(ntiles, t * seg), e = min(ntiles, a + seg);
  unsigned sum = 0;
  for (int i = a; i < e; ++i) sum += src[i];
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {  // inclusive Hillis-Steele
    const unsigned v = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  unsigned run = base + s[t] - sum;
  for (int i = a; i < e; ++i) {
    const unsigned v = src[i];
    src[i] = run;
    run += v;
  }
  if (staged) {
    __syncthreads();
    for (int i = t; i < ntiles; i += kThreads) row[i] = rowl[i];
  }
}

// Stable scatter of one pass.  payload_in == nullptr means the identity.  A
// trivial pass copies keys and payload unchanged.  The keys of a tile are
// ranked as k_seg_sort ranks a segment's: every wave counts the digits of
// all R rounds at once (one barrier for the whole tile), thread t turns
// digit t's counts into an exclusive prefix over (round, wave), which is
// tile order, and every key gets its slot in the tile's digit order before
// any is staged.  The counts are 16-bit (a tile holds 256 R <= 4096 keys)
// and share LDS with the staged keys, which are written only after every
// slot is known: 50 KB of LDS, three workgroups per CU.
//
// Persistent: XCD x (blockIdx & 7) owns the contiguous tiles [x per, (x+1)
// per), and its workgroups take them in turn, so the tiles in flight on one
// XCD are neighbours: the runs that neighbouring tiles write for one digit
// are neighbours in kout, and the cache lines they share are completed in
// that XCD's L2.  A workgroup loads the keys and payload of its next tile
// into registers once the current tile's are staged in LDS, so the loads
// overlap the current tile's writes.  The grid is a multiple of 8.
template <int R>
__global__ __launch_bounds__(kThreads, 3) void k_scatter(const uint64_t* __restrict__ kin,
                                                      const unsigned* __restrict__ pin,
                                                      uint64_t* __restrict__ kout,
                                                      unsigned* __restrict__ pout, int64_t n,
                                                      int pass, int ntiles,
                                                      const unsigned* __restrict__ offsets,
                                                      const unsigned* __restrict__ hist_pass) {
  constexpr int kRounds = R, kTile = kThreads * R;
  static_assert(kTile <= 65535, "16-bit tile ranks");
  static_assert(kThreads == 256, "one thread per digit");
  const int per = (ntiles + 7) >> 3, xcd = blockIdx.x & 7, nwx = gridDim.x >> 3;
  const int tend = min(ntiles, (xcd + 1) * per);
  int tile = xcd * per + (int)(blockIdx.x >> 3);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (pass_trivial(hist_pass, n)) {
    for (; tile < tend; tile += nwx) {
      const int64_t base = (int64_t)tile * kTile;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int64_t i = base + r * kThreads + t;
        if (i < n) {
          kout[i] = kin[i];
          pout[i] = pin ? pin[i] : (unsigned)i;
        }
      }
    }
    return;
  }
  // cnt[r][w][d] (ranking) and the staged keys (placement) share this space
  __shared__ __align__(16) uint64_t sk[kTile];
  static_assert(sizeof(unsigned short) * kRounds * kWaves * 256 <= sizeof(uint64_t) * kTile,
                "the rank counts fit where the keys are staged");
  unsigned short(*cnt)[kWaves][256] = reinterpret_cast<unsigned short(*)[kWaves][256]>(sk);
  __shared__ unsigned sp[kTile];
  __shared__ unsigned wsum[kWaves];
  __shared__ unsigned start[256];  // first LDS slot of each digit (exclusive scan)
  __shared__ unsigned goff[256];   // where this tile's run of each digit starts in kout
  // the keys, payload and run offsets of the tile to rank next
  uint64_t key[kRounds];
  unsigned pay[kRounds], on = 0;
  auto load = [&](int tl) {
    const int64_t base = (int64_t)tl * kTile;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int64_t i = base + r * kThreads + t;
      const bool ok = i < n;
      key[r] = ok ? kin[i] : ~0ull;
      pay[r] = ok ? (pin ? pin[i] : (unsigned)i) : 0u;
    }
    on = offsets[(size_t)t * ntiles + tl];
  };
  if (tile < tend) load(tile);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (; tile < tend; tile += nwx) {  // uniform per workgroup
    const int64_t base = (int64_t)tile * kTile;
    unsigned dig[kRounds], slot[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
      dig[r] = base + r * kThreads + t < n ? digit_of(key[r], pass) : 256u;  // 256: padding, not ranked
    goff[t] = on;
    for (int i = t; i < kRounds * kWaves * 256 / 8; i += kThreads)
      reinterpret_cast<uint4*>(sk)[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    // rank within the wave: lanes of this wave holding the same digit
    // (padding lanes match nobody); the first of them records the count
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const unsigned d = dig[r];
      uint64_t peers = __ballot(d < 256u);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
      }
      if (d >= 256u) peers = 0;
      slot[r] = (unsigned)__popcll(peers & lt);
      if (d < 256u && slot[r] == 0) cnt[r][w][d] = (unsigned short)__popcll(peers);
    }
    __syncthreads();
    // thread t owns digit t: exclusive prefix over (round, wave), then the
    // digit's start from a scan of the totals over the digits
    unsigned tot = 0;
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) {
        const unsigned v = cnt[r][ww][t];
        cnt[r][ww][t] = (unsigned short)tot;
        tot += v;
      }
    unsigned incl = tot;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned u = __shfl_up(incl, off, 64);
      if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    unsigned pre = 0;
    for (int ww = 0; ww < w; ++ww) pre += wsum[ww];
    start[t] = pre + incl - tot;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
      if (dig[r] < 256u) slot[r] += start[dig[r]] + cnt[r][w][dig[r]];
    __syncthreads();  // every slot read its counts: the space now takes the keys
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      if (dig[r] < 256u) {
        sk[slot[r]] = key[r];
        sp[slot[r]] = pay[r];
      }
    }
    __syncthreads();
    if (tile + nwx < tend) load(tile + nwx);  // this tile's keys are in LDS now
    // out in runs that are contiguous in memory
    const int valid = (int)min<int64_t>(kTile, n - base);
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int q = r * kThreads + t;
      if (q < valid) {
        const uint64_t k = sk[q];
        const unsigned d = digit_of(k, pass);
        const size_t dst = (size_t)goff[d] + (unsigned)(q - start[d]);
        kout[dst] = k;
        pout[dst] = sp[q];
      }
    }
    __syncthreads();  // the next tile's counts overwrite sk and goff
  }
}

// ---- sorts of up to kSmallSort keys: segments sorted in LDS, then merged.
// kSegThreads threads (kSegWaves waves) per segment of kSeg keys, kSegR keys
// a thread: sixteen waves a CU cover the latency of each pass's LDS and
// barrier chain, which one wave per SIMD left exposed.
constexpr int kSegThreads = 1024, kSegR = 2, kSeg = kSegThreads * kSegR;  // keys per segment
constexpr int kSegWaves = kSegThreads / 64;
constexpr int kSegParts = kSegThreads / 256;  // threads per digit in the scan
constexpr int kSegEnt = kSegR * kSegWaves;    // (round, wave) entries per digit
static_assert(kSegEnt % kSegParts == 0, "the scan splits a digit's entries evenly");
constexpr int kMergeThreads = 256, kMergeItems = 2;  // outputs per thread of a merge round
constexpr int kMergeTile = kMergeThreads * kMergeItems;  // outputs per workgroup
static_assert((2 * kSeg) % kMergeTile == 0, "a merge workgroup stays inside one pair of runs");

// Stable sort of one segment of kSeg keys (the doubles, turned into ordered
// integer keys as k_transform_hist does), with their original indices as
// payload, entirely in LDS: eight 8-bit LSD passes, each ranking keys stably
// as k_scatter does (segment order = (round, wave, lane)); a pass whose
// digit is the same for every key of the segment is the identity and is
// skipped.
__global__ __launch_bounds__(kSegThreads) void k_seg_sort(const double* __restrict__ in, int64_t n,
                                                          uint64_t* __restrict__ kout,
                                                          unsigned* __restrict__ pout) {
  __shared__ uint64_t sk[kSeg];
  __shared__ unsigned sp[kSeg];
  // cnt[e][d], e = round * kSegWaves + wave: keys of digit d in that (round,
  // wave); then, in place, the number of keys of digit d before it in
  // segment order
  __shared__ __align__(16) unsigned cnt[kSegEnt][256];
  __shared__ unsigned ptot[kSegParts][256];  // keys of digit d in each part's entries
  __shared__ unsigned wsum[4];
  __shared__ unsigned start[256];  // first slot of each digit (exclusive scan)
  __shared__ uint64_t red[2][kSegWaves];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * kSeg;
  const int valid = (int)min<int64_t>(kSeg, n - base);
  uint64_t kand = ~0ull, kor = 0ull;  // bits shared by all keys of the segment
  for (int q = t; q < kSeg; q += kSegThreads) {
    const uint64_t k = q < valid ? key_of(in[base + q]) : ~0ull;
    sk[q] = k;
    sp[q] = (unsigned)(base + q);
    if (q < valid) {
      kand &= k;
      kor |= k;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    kand &= __shfl_xor(kand, off, 64);
    kor |= __shfl_xor(kor, off, 64);
  }
  if (lane == 0) {
    red[0][w] = kand;
    red[1][w] = kor;
  }
  __syncthreads();
  uint64_t varying = 0;  // bits that differ between two keys of the segment
  {
    uint64_t a = ~0ull, o = 0ull;
    for (int ww = 0; ww < kSegWaves; ++ww) {
      a &= red[0][ww];
      o |= red[1][ww];
    }
    varying = a ^ o;
  }
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the scan: thread t takes digit sd and entries [se0, se0 + kSegEnt/kSegParts)
  const int sd = t & 255, sj = t >> 8, se0 = sj * (kSegEnt / kSegParts);
  for (int pass = 0; pass < 8; ++pass) {
    if (((varying >> (8 * pass)) & 255u) == 0) continue;  // block-uniform
    for (int i = t; i < kSegEnt * 64; i += kSegThreads)
      reinterpret_cast<uint4*>(&cnt[0][0])[i] = make_uint4(0u, 0u, 0u, 0u);
    uint64_t key[kSegR];
    unsigned pay[kSegR], dig[kSegR], before[kSegR];
#pragma unroll
    for (int r = 0; r < kSegR; ++r) {
      const int q = r * kSegThreads + t;
      key[r] = sk[q];
      pay[r] = sp[q];
      dig[r] = q < valid ? digit_of(key[r], pass) : 256u;  // 256: padding, not ranked
    }
    __syncthreads();
    // rank within the wave: lanes of this wave holding the same digit
    // (padding lanes match nobody); the first of them records the count
#pragma unroll
    for (int r = 0; r < kSegR; ++r) {
      const unsigned d = dig[r];
      uint64_t peers = __ballot(d < 256u);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
      }
      if (d >= 256u) peers = 0;
      before[r] = (unsigned)__popcll(peers & lt);
      if (d < 256u && before[r] == 0) cnt[r * kSegWaves + w][d] = (unsigned)__popcll(peers);
    }
    __syncthreads();
    // per digit, kSegParts threads each sum a stretch of its entries ...
    unsigned part = 0;
#pragma unroll
    for (int e = 0; e < kSegEnt / kSegParts; ++e) part += cnt[se0 + e][sd];
    ptot[sj][sd] = part;
    __syncthreads();
    // ... then turn their stretch into exclusive prefixes in segment order;
    // the first 256 threads also scan the digit totals
    unsigned run = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < kSegParts; ++j) {
      const unsigned v = ptot[j][sd];
      if (j < sj) run += v;
      tot += v;
    }
#pragma unroll
    for (int e = 0; e < kSegEnt / kSegParts; ++e) {
      const unsigned v = cnt[se0 + e][sd];
      cnt[se0 + e][sd] = run;
      run += v;
    }
    unsigned incl = tot;
    if (sj == 0) {
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const unsigned u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
      }
      if (lane == 63) wsum[w] = incl;
    }
    __syncthreads();
    if (sj == 0) {
      unsigned pre = 0;
      for (int ww = 0; ww < w; ++ww) pre += wsum[ww];
      start[sd] = pre + incl - tot;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegR; ++r) {
      if (dig[r] < 256u) {
        const unsigned q = start[dig[r]] + cnt[r * kSegWaves + w][dig[r]] + before[r];
        sk[q] = key[r];
        sp[q] = pay[r];
      }
    }
    __syncthreads();
  }
  for (int q = t; q < valid; q += kSegThreads) {
    kout[base + q] = sk[q];
    pout[base + q] = sp[q];
  }
}

// Merge path: the number of elements of the left run A (length la) among
// the first d outputs of the stable merge of A and B (length lb), i.e. the
// first i with A[i] > B[d - 1 - i].  One wavefront searches together: 64
// probes per step narrow the range 64-fold, so a run of 2^20 needs four
// dependent rounds of reads instead of twenty.
__device__ __forceinline__ int64_t merge_split_wave(const uint64_t* __restrict__ A, int64_t la,
                                                    const uint64_t* __restrict__ B, int64_t lb, int64_t d,
                                                    int lane) {
  int64_t lo = max<int64_t>(0, d - lb), hi = min<int64_t>(d, la);
  while (hi > lo) {  // uniform per wave
    const int64_t span = hi - lo;
    const int64_t p = span <= 64 ? lo + lane : lo + span * lane / 64;
    const bool in = p < hi && A[p] <= B[d - 1 - p];
    const int cnt = __popcll(__ballot(in));  // the probes that hold form a prefix
    if (span <= 64) return lo + cnt;
    const int64_t p_last = lo + span * (cnt - 1) / 64, p_first_out = lo + span * cnt / 64;
    const int64_t nlo = cnt == 0 ? lo : p_last + 1;
    const int64_t nhi = cnt == 64 ? hi : p_first_out;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// One merge round: adjacent sorted runs of length L (the last ones may be
// shorter) become runs of 2L.  A workgroup writes kMergeTile consecutive
// outputs, all inside one pair of runs (2L is a multiple of kMergeTile):
// two wavefronts find where its first and last outputs split the pair, the
// workgroup copies those stretches of both runs to LDS, and each thread
// merges kMergeItems outputs there after its own merge-path search; the
// merged tile goes back through LDS so that its stores are contiguous.
// Stable: on equal keys the left run's element comes first.
__global__ __launch_bounds__(kMergeThreads) void k_merge(const uint64_t* __restrict__ kin,
                                                         const unsigned* __restrict__ pin,
                                                         uint64_t* __restrict__ kout,
                                                         unsigned* __restrict__ pout, int64_t n, int64_t L) {
  __shared__ uint64_t sk[kMergeTile];
  __shared__ unsigned sp[kMergeTile];
  __shared__ int64_t split[2];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t o0 = (int64_t)blockIdx.x * kMergeTile;
  const int64_t q0 = o0 / (2 * L) * (2 * L);
  const int64_t a0 = q0, a1 = min(q0 + L, n), b0 = a1, b1 = min(q0 + 2 * L, n);
  const int64_t la = a1 - a0, lb = b1 - b0;
  const int64_t d0 = o0 - q0, d1 = min<int64_t>(d0 + kMergeTile, la + lb);
  if (w < 2) {
    const int64_t sp_ = merge_split_wave(kin + a0, la, kin + b0, lb, w == 0 ? d0 : d1, lane);
    if (lane == 0) split[w] = sp_;
  }
  __syncthreads();
  const int64_t i0 = split[0], i1 = split[1];
  const int na = (int)(i1 - i0), nb = (int)((d1 - i1) - (d0 - i0));
  for (int q = t; q < na + nb; q += kMergeThreads) {
    const int64_t src = q < na ? a0 + i0 + q : b0 + (d0 - i0) + (q - na);
    sk[q] = kin[src];
    sp[q] = pin[src];
  }
  __syncthreads();
  const int e0 = t * kMergeItems, ne = (int)(d1 - d0);
  uint64_t ok[kMergeItems];
  unsigned op[kMergeItems];
  if (e0 < ne) {
    int lo = max(0, e0 - nb), hi = min(e0, na);
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sk[mid] <= sk[na + e0 - 1 - mid]) lo = mid + 1;
      else hi = mid;
    }
    int i = lo, j = na + (e0 - lo);
    const int jend = na + nb;
#pragma unroll
    for (int e = 0; e < kMergeItems; ++e) {
      if (e0 + e < ne) {
        const bool take_a = j >= jend || (i < na && sk[i] <= sk[j]);
        const int src = take_a ? i++ : j++;
        ok[e] = sk[src];
        op[e] = sp[src];
      }
    }
  }
  // through LDS again, so that the stores to global memory are contiguous
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kMergeItems; ++e)
    if (e0 + e < ne) {
      sk[e0 + e] = ok[e];
      sp[e0 + e] = op[e];
    }
  __syncthreads();
  for (int q = t; q < ne; q += kMergeThreads) {
    kout[o0 + q] = sk[q];
    pout[o0 + q] = sp[q];
  }
}

// The double whose ordered key is k, for every key that determines it: all
// but the key of +-0.0 (both are 0x8000000000000000) and that of NaN (~0).
__device__ __forceinline__ double double_of(uint64_t k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

// rank[perm[i]] = i and sorted[i] = keys[perm[i]].  The sorted values come
// from the sorted ordered keys skeys (read in order); only a key that does
// not determine its double (a zero, whose sign it dropped, or a NaN, whose
// payload it dropped) is gathered from the original keys.
// Only ranks of the pairs p in [plo, phi) are written (see sort_f64).
__global__ void k_rank_and_gather(const unsigned* __restrict__ perm, const uint64_t* __restrict__ skeys,
                                  const double* __restrict__ keys, int64_t n, unsigned* __restrict__ rank,
                                  double* __restrict__ sorted, unsigned plo, unsigned phi) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned p = perm[i];
    if (rank && p - plo < phi - plo) rank[p] = (unsigned)i;
    if (sorted) {
      const uint64_t k = skeys[i];
      sorted[i] = (k == ~0ull || k == 0x8000000000000000ull) ? keys[p] : double_of(k);
    }
  }
}

// upper[i] = #{non-NaN keys <= bins[i]} of an ascending array (NaNs last).
__global__ void k_upper_bound(const double* __restrict__ s, int64_t n, const double* __restrict__ bins,
                              int nb, unsigned* __restrict__ upper) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  // number of non-NaN keys
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    double v = s[mid];
    if (v != v) hi = mid;
    else lo = mid + 1;
  }
  const int64_t nvalid = lo;
  const double b = bins[i];
  lo = 0;
  hi = nvalid;
  while (lo < hi) {  // first index with b < s[idx]
    int64_t mid = (lo + hi) >> 1;
    if (b < s[mid]) hi = mid;
    else lo = mid + 1;
  }
  upper[i] = (unsigned)lo;
}

// First index in [lo, hi) whose key satisfies pred (pred is false, then
// true along the range), or hi.  All 64 lanes of a wavefront call it
// together: each round probes 64 evenly spaced keys at once and keeps the
// gap between the last false and the first true probe, so a range of n keys
// takes about log64(n) dependent reads instead of log2(n).
template <class Pred>
__device__ int64_t wave_partition_point(const double* __restrict__ s, int64_t lo, int64_t hi, Pred pred) {
  const int l = threadIdx.x & 63;
  while (hi - lo > 64) {
    const int64_t step = (hi - lo + 63) / 64;
    const int64_t p = lo + (int64_t)l * step;
    const bool valid = p < hi;
    const unsigned long long bt = __ballot(valid && pred(s[p]));
    if (bt == 0) {
      const int last = 63 - __builtin_clzll(__ballot(valid));
      lo += (int64_t)last * step + 1;
    } else {
      const int f = __builtin_ctzll(bt);
      if (f == 0) return lo;
      hi = lo + (int64_t)f * step;
      lo += (int64_t)(f - 1) * step + 1;
    }
  }
  const int64_t p = lo + l;
  const unsigned long long bt = __ballot(p < hi && pred(s[p]));
  return bt ? lo + __builtin_ctzll(bt) : hi;
}

// The bin table of an ascending array (NaNs last) in one launch, one
// wavefront per bin: upper[i] = #{non-NaN keys <= bins[i]}, lower[i] =
// upper[i-1] (0 for i = 0), npb[i] = upper[i] - lower[i] -- the values
// of setup_integral_GPU's upper_bound and adjacent_difference
// (gpu_functions.cu:386-397); k_upper_bound gives the same upper[i].
__global__ __launch_bounds__(64) void k_bin_table_waves(const double* __restrict__ s, int64_t n,
                                                        const double* __restrict__ bins, int nb,
                                                        unsigned* __restrict__ upper,
                                                        unsigned* __restrict__ lower,
                                                        unsigned* __restrict__ npb) {
  const int i = blockIdx.x;
  if (i >= nb) return;
  const int64_t nvalid = wave_partition_point(s, 0, n, [](double v) { return v != v; });
  auto ub = [&](double b) {
    return (unsigned)wave_partition_point(s, 0, nvalid, [b](double v) { return b < v; });
  };
  const unsigned up = ub(bins[i]);
  const unsigned lo = i ? ub(bins[i - 1]) : 0u;
  if (threadIdx.x == 0) {
    upper[i] = up;
    lower[i] = lo;
    npb[i] = up - lo;
  }
}

template <int R>
// Returns the buffer that holds the sorted keys.
const uint64_t* sort_passes(mlc_ctx* ctx, uint64_t* k0, uint64_t* k1, unsigned* pbuf, unsigned* perm,
                            unsigned* cnt, const unsigned* hist, int64_t n) {
  constexpr int kTile = kThreads * R;
  const int ntiles = (int)((n + kTile - 1) / kTile);
  const int staged = ntiles <= kScanStageTiles;
  const size_t scan_lds = staged ? sizeof(unsigned) * ntiles : 0;
  if (scan_lds > 64 * 1024)
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scan_tiles),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)scan_lds),
              "hipFuncSetAttribute(scan LDS)");
  // persistent tile counts: eight workgroups per CU, a multiple of 8
  const int hist_grid = (int)std::min<int64_t>(cdiv(ntiles, 8) * 8, cdiv(fill_grid(ctx, 8), 8) * 8);
  // persistent scatter: the workgroups the CUs hold at once, a multiple of 8
  const int scatter_grid =
      (int)std::min<int64_t>(cdiv(ntiles, 8) * 8,
                             cdiv(fill_grid(ctx, resident_blocks(reinterpret_cast<const void*>(&k_scatter<R>),
                                                                 kThreads)), 8) * 8);
  // eight passes (trivial ones copy): ping-pong so that the last one writes perm
  unsigned* pa = pbuf;  // payload written by passes 0, 2, 4, 6
  unsigned* pb = perm;  // payload written by passes 1, 3, 5, 7
  const uint64_t* kin = k0;
  uint64_t* kout = k1;
  const unsigned* pin = nullptr;
  for (int p = 0; p < 8; ++p) {
    unsigned* pout = (p % 2 == 0) ? pa : pb;
    const unsigned* hp = hist + p * 256;
    MLC_LAUNCH(ctx, "sort_tile_hist", k_tile_hist<R>, dim3(hist_grid), dim3(kThreads), 0, kin, n, p, ntiles,
               cnt, hp);
    MLC_LAUNCH(ctx, "sort_scan_tiles", k_scan_tiles, dim3(256), dim3(kThreads), scan_lds, cnt, ntiles, hp, n,
               staged);
    MLC_LAUNCH(ctx, "sort_scatter", k_scatter<R>, dim3(scatter_grid), dim3(kThreads), 0, kin, pin, kout, pout,
               n, p, ntiles, cnt, hp);
    const uint64_t* nk = kout;
    kout = (kout == k1) ? k0 : k1;
    kin = nk;
    pin = pout;
  }
  return kin;
}

}  // namespace

namespace mlc {

// Stable argsort into d_perm (sorted position -> original index).  Never
// synchronises with the host.
void sort_f64(mlc_ctx* ctx, const double* keys, int64_t n, unsigned* perm, unsigned* rank,
              double* sorted) {
  if (n <= 0) return;
  MLC_REQUIRE(n < ((int64_t)1 << 32), "sort: more than 2^32-1 keys");
  const bool small = n <= kSmallSort;
  const int64_t tile = kThreads * kScatterRounds;
  const int64_t ntiles = small ? 1 : (n + tile - 1) / tile;
  Workspace ws(ctx);
  int ik0 = ws.add(sizeof(uint64_t) * n), ik1 = ws.add(sizeof(uint64_t) * n);
  int ip1 = ws.add(sizeof(unsigned) * n);
  int ih = ws.add(sizeof(unsigned) * 8 * 256);
  int ic = ws.add(sizeof(unsigned) * 256 * (size_t)ntiles);
  ws.commit();
  uint64_t *k0 = ws.get<uint64_t>(ik0), *k1 = ws.get<uint64_t>(ik1);
  unsigned *pbuf = ws.get<unsigned>(ip1), *hist = ws.get<unsigned>(ih), *cnt = ws.get<unsigned>(ic);
  const uint64_t* skeys = nullptr;  // the sorted ordered keys
  if (small) {
    // segments, then merge rounds; the segment sort writes the payload where
    // an even number of rounds leaves it in perm
    int rounds = 0;
    for (int64_t L = kSeg; L < n; L *= 2) ++rounds;
    unsigned* p0 = rounds % 2 == 0 ? perm : pbuf;
    MLC_LAUNCH(ctx, "sort_segments", k_seg_sort, dim3(cdiv(n, kSeg)), dim3(kSegThreads), 0, keys, n, k1, p0);
    uint64_t *kin = k1, *kout = k0;
    unsigned *pin = p0, *pout = p0 == perm ? pbuf : perm;
    for (int64_t L = kSeg; L < n; L *= 2) {
      MLC_LAUNCH(ctx, "sort_merge", k_merge, dim3(cdiv(n, kMergeTile)), dim3(kMergeThreads), 0, kin, pin,
                 kout, pout, n, L);
      std::swap(kin, kout);
      std::swap(pin, pout);
    }
    skeys = kin;
  } else {
    // a fill kernel, not a memset node, so that a captured graph replays it
    MLC_LAUNCH(ctx, "sort_zero_hist", k_zero_u32, dim3(8), dim3(256), 0, hist, 8 * 256);
    const int g = (int)std::min<int64_t>(cdiv(n, kHistThreads * 4), fill_grid(ctx, 2));
    MLC_LAUNCH(ctx, "sort_transform_hist", k_transform_hist, dim3(g), dim3(kHistThreads), 0, keys, k0, n,
               hist);
    skeys = sort_passes<kScatterRounds>(ctx, k0, k1, pbuf, perm, cnt, hist, n);
  }
  if (rank || sorted) {
    // The rank scatter writes 4 bytes to random places.  Over an array larger
    // than the last-level cache every such write reaches HBM as a partial
    // line; in windows of kRankWindow ranks, each launch reading all of perm
    // but writing only into its window, the partial lines of a window meet
    // in the cache first.
    const int64_t win = rank ? kRankWindow : n;
    for (int64_t lo = 0; lo < n; lo += win)
      MLC_LAUNCH(ctx, "sort_rank_gather", k_rank_and_gather,
                 dim3(std::min<int64_t>(cdiv(n, 256), fill_grid(ctx, 16))), dim3(256), 0, perm, skeys, keys, n,
                 rank, lo == 0 ? sorted : nullptr, (unsigned)lo, (unsigned)std::min<int64_t>(n, lo + win));
  }
}

// The bin table on the device: d_lower / d_npb from the upper bounds, no
// host round trip.
void bin_edges_device(mlc_ctx* ctx, const double* sorted, int64_t n, const double* d_bins, int Nbins,
                      unsigned* d_upper, unsigned* d_lower, unsigned* d_npb) {
  MLC_LAUNCH(ctx, "bin_table", k_bin_table_waves, dim3(Nbins), dim3(64), 0, sorted, n, d_bins, Nbins,
             d_upper, d_lower, d_npb);
}

void bin_edges(mlc_ctx* ctx, const double* sorted, int64_t n, const double* bins, int Nbins,
               unsigned* lower_ind, unsigned* n_per_bin) {
  Workspace ws(ctx);
  int ib = ws.add(sizeof(double) * Nbins), iu = ws.add(sizeof(unsigned) * Nbins);
  ws.commit();
  double* db = ws.get<double>(ib);
  unsigned* du = ws.get<unsigned>(iu);
  h2d(ctx, db, bins, sizeof(double) * Nbins);
  MLC_LAUNCH(ctx, "bin_upper_bound", k_upper_bound, dim3(cdiv(Nbins, 64)), dim3(64), 0, sorted, n,
             db, Nbins, du);
  std::vector<unsigned> up(Nbins);
  d2h(ctx, up.data(), du, sizeof(unsigned) * Nbins);
  sync(ctx);
  lower_ind[0] = 0;
  for (int i = 1; i < Nbins; ++i) lower_ind[i] = up[i - 1];
  n_per_bin[0] = up[0];
  for (int i = 1; i < Nbins; ++i) n_per_bin[i] = up[i] - up[i - 1];
}

}  // namespace mlc

extern "C" {

int mlc_sort_f64(mlc_ctx* ctx, const double* d_keys, int64_t n, uint32_t* d_perm, uint32_t* d_rank,
                 double* d_sorted) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(n >= 0, "mlc_sort_f64: negative n");
    MLC_REQUIRE(n == 0 || (d_keys && d_perm), "mlc_sort_f64: null pointer");
    sort_f64(ctx, d_keys, n, d_perm, d_rank, d_sorted);
  });
}

int mlc_bin_edges(mlc_ctx* ctx, const double* d_sorted, int64_t n, const double* bins, int Nbins,
                  uint32_t* lower_ind, uint32_t* n_per_bin) {
  return guarded(ctx, [&] {
    MLC_REQUIRE(Nbins > 0 && bins && lower_ind && n_per_bin, "mlc_bin_edges: bad bins");
    MLC_REQUIRE(n >= 0 && (n == 0 || d_sorted), "mlc_bin_edges: bad keys");
    bin_edges(ctx, d_sorted, n, bins, Nbins, lower_ind, n_per_bin);
  });
}

}  // extern "C"
