// Device kernels of the MI355X LBVH gravity step (gfx950, wave64).
//
// Every kernel is segment-aware: it iterates over a table of Tile descriptors
// ({segment, first element, count}) with a grid-stride loop, so one launch covers
// one system or many uneven systems (batched/served forms) alike.
#include "glx_kernels.h"

#include <math.h>

namespace glx {

// ----------------------------------------------------------------------------------
// helpers
// ----------------------------------------------------------------------------------

// Order-preserving map f64 -> u64 (non-NaN inputs): unsigned compare of the
// encodings equals f64 compare, so min/max reduce with integer atomics exactly.
__device__ __forceinline__ unsigned long long ord_enc(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord_dec(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_minf(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide exclusive scan of one u32 per thread, any blockDim that is a multiple of
// 64; sh_wave must hold blockDim / 64 entries (one per wave).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh_wave) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) sh_wave[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int k = 0; k < w; ++k) pre += sh_wave[k];
  __syncthreads();
  return pre + inc - v;
}

// Reduce a per-thread bbox over the block and fold it into the segment's atomic
// accumulators. min/max are exact and order-independent, so the result is
// bit-identical to a serial fold whatever the arrival order.
__device__ void block_commit_bbox(double lo[3], double hi[3], unsigned long long* bbox,
                                  double* sh /* 6 * waves */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int a = 0; a < 3; ++a) {
    lo[a] = wave_min(lo[a]);
    hi[a] = wave_max(hi[a]);
  }
  if (lane == 0)
    for (int a = 0; a < 3; ++a) {
      sh[w * 6 + a] = lo[a];
      sh[w * 6 + 3 + a] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    double r = sh[a];
    for (int k = 1; k < kBlock / 64; ++k)
      r = (a < 3) ? fmin(r, sh[k * 6 + a]) : fmax(r, sh[k * 6 + a]);
    // a NaN-only tile contributes the neutral element (f64::min/max ignore NaN)
    if (!isnan(r)) {
      if (a < 3)
        atomicMin(&bbox[a], ord_enc(r));
      else
        atomicMax(&bbox[a], ord_enc(r));
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void bbox_neutral(unsigned long long* bbox, int a) {
  bbox[a] = (a < 3) ? 0xFFFFFFFFFFFFFFFFull : 0ull;
}

// solvers/src/lbvh.rs:37-50
__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
// the same spread over 21 bits (63-bit codes)
__device__ __forceinline__ uint64_t expand21(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

template <class K>
struct KeyTraits;
template <>
struct KeyTraits<uint32_t> {
  static constexpr int cell_bits = 10;  // solvers/src/lbvh.rs:32
  static constexpr int tie_base = 32;
  __device__ static uint32_t code(uint32_t x, uint32_t y, uint32_t z) {
    return expand10(x) | (expand10(y) << 1) | (expand10(z) << 2);
  }
  __device__ static int clz(uint32_t v) { return __clz((int)v); }
};
template <>
struct KeyTraits<uint64_t> {
  static constexpr int cell_bits = 21;  // the 63-bit refinement of solvers/src/lbvh.rs:29-31
  static constexpr int tie_base = 64;
  __device__ static uint64_t code(uint32_t x, uint32_t y, uint32_t z) {
    return expand21(x) | (expand21(y) << 1) | (expand21(z) << 2);
  }
  __device__ static int clz(uint64_t v) { return __clzll((long long)v); }
};

// solvers/src/lbvh.rs:110-120: the padded root cube of the segment's bbox, in f64
// with the reference's operation order (no contraction). origin = cube centre.
struct Bounds {
  double bmin[3], size, scale, origin[3];
};
__device__ Bounds seg_bounds(const unsigned long long* bbox, int cell_bits) {
#pragma clang fp contract(off)
  Bounds b;
  double lo[3], hi[3], ext[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = ord_dec(bbox[a]);
    hi[a] = ord_dec(bbox[3 + a]);
  }
  for (int a = 0; a < 3; ++a) {
    b.origin[a] = (lo[a] + hi[a]) * 0.5;
    ext[a] = hi[a] - lo[a];
  }
  const double maxe = fmax(ext[0], fmax(ext[1], ext[2]));
  const double half = fmax(0.5 * maxe, 1e-12) * (1.0 + 1e-9);
  for (int a = 0; a < 3; ++a) b.bmin[a] = b.origin[a] - half;
  b.size = 2.0 * half;
  b.scale = (double)(1u << cell_bits) / b.size;
  return b;
}

// solvers/src/lbvh.rs:56-64: floor, clamp below at 0 (NaN -> 0), saturate at the top cell.
__device__ __forceinline__ uint32_t quant_lane(double v, double top) {
  const double f = fmax(floor(v), 0.0);
  return f >= top ? (uint32_t)top : (uint32_t)f;
}

// A pointer from a SegDesc as a global-memory pointer: its loads and stores become
// global_load / global_store (counted by vmcnt only) instead of flat ones, which also
// hold up every LDS wait.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gbl(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gblw(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// MI355X deals the blocks of a launch round-robin over its 8 XCDs (block b runs on XCD
// b % 8), each with its own L2. Walk groups that are neighbours in Morton order read
// nearly the same nodes, so the walks renumber blocks: in every span of 8 * kXcdRun
// blocks, the kXcdRun blocks an XCD receives take kXcdRun consecutive tiles. Neighbours
// then share an L2 while all XCDs still advance through the tiles together (a whole-range
// split per XCD leaves XCDs with unequal work). Launches use a grid that is a multiple of
// the span (xcd_grid); any other grid keeps the hardware order. Runs of 256 walk tiles
// (16,384 sorted bodies; an XCD holds about two runs at a time) measured fastest on the
// 512-system ensemble: walk -7% against runs of 8, -4.5% against 64, +0.6% at 512.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdRun = 256;
constexpr uint32_t kXcdSpan = kXcds * kXcdRun;

template <uint32_t RUN = kXcdRun>
__device__ __forceinline__ uint32_t xcd_block() {
  constexpr uint32_t span = kXcds * RUN;
  const uint32_t g = gridDim.x, b = blockIdx.x;
  if (g % span) return b;
  const uint32_t w = b % span;
  return (b - w) + (w % kXcds) * RUN + w / kXcds;
}

// The build kernels over sort tiles (2048 keys) take runs of 32 tiles per XCD, so that a
// system of up to 65,536 bodies is gathered, sorted and linked on one XCD and its random
// reads (positions by sorted index, codes in the binary searches) and scattered writes
// stay in that XCD's L2.
constexpr uint32_t kXcdBuildRun = 32;

// The per-target walk keeps runs of 8 tiles: its cost per target varies far more across a
// system (dense cores against halos), and long runs left whole cores to one XCD (2^20
// collision, reference build: 1.8x slower at 256).
constexpr uint32_t kXcdLaneRun = 8;

// ----------------------------------------------------------------------------------
// 1. bounding box
// ----------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_bbox(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  __shared__ double sh[6 * (kBlock / 64)];
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const size_t i = tl.start + e;
      for (int a = 0; a < 3; ++a) {
        const double p = s.x[3 * i + a];
        lo[a] = fmin(lo[a], p);
        hi[a] = fmax(hi[a], p);
      }
    }
    block_commit_bbox(lo, hi, s.bbox, sh);
  }
}

// Opening half-kick with a(x_n), f64 drift (f32 on the all-f32 path), and the bbox of
// the drifted positions
// (core/src/integrator.rs:59-67; gpu/src/gpu_resident.rs:79-159).
__global__ void __launch_bounds__(kBlock)
k_kick_drift_bbox(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles,
                  uint32_t ntiles) {
#pragma clang fp contract(off)
  __shared__ double sh[6 * (kBlock / 64)];
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const float h = s.half_dt;
    const double dt = s.dt;
    auto* sv = gblw(s.v);
    auto* sx = gblw(s.x);
    const auto* sa = gbl(s.acc);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (s.precision == kPrecAbs32) {  // the all-f32 path drifts in f32 (tile-uniform)
      const float dt32 = (float)dt;
      for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
        const size_t i = tl.start + e;
        for (int a = 0; a < 3; ++a) {
          const float v = sv[3 * i + a] + sa[3 * i + a] * h;
          sv[3 * i + a] = v;
          const double x = (double)((float)sx[3 * i + a] + v * dt32);
          sx[3 * i + a] = x;
          lo[a] = fmin(lo[a], x);
          hi[a] = fmax(hi[a], x);
        }
      }
    } else {
      for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
        const size_t i = tl.start + e;
        for (int a = 0; a < 3; ++a) {
          const float v = sv[3 * i + a] + sa[3 * i + a] * h;
          sv[3 * i + a] = v;
          const double x = sx[3 * i + a] + (double)v * dt;
          sx[3 * i + a] = x;
          lo[a] = fmin(lo[a], x);
          hi[a] = fmax(hi[a], x);
        }
      }
    }
    block_commit_bbox(lo, hi, s.bbox, sh);
  }
}

// ----------------------------------------------------------------------------------
// 2. Morton encode (bit-exact with solvers/src/lbvh.rs:105-129; 63-bit variant with
//    2^21 cells per axis)
// ----------------------------------------------------------------------------------

template <class K>
__global__ void __launch_bounds__(kBlock)
k_encode(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
         uint32_t* __restrict__ lanes_out, int dst) {
#pragma clang fp contract(off)
  using T = KeyTraits<K>;
  const double top = (double)((1u << T::cell_bits) - 1u);
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const Bounds b = seg_bounds(s.bbox, T::cell_bits);
    if (tl.start == 0 && threadIdx.x == 0) {
      s.bounds[0] = b.bmin[0];
      s.bounds[1] = b.bmin[1];
      s.bounds[2] = b.bmin[2];
      s.bounds[3] = b.size;
      s.bounds[4] = b.scale;
      // walk coordinates are relative to the cube centre, or absolute on the all-f32 path
      const bool abs32 = s.precision == kPrecAbs32;
      s.bounds[5] = abs32 ? 0.0 : b.origin[0];
      s.bounds[6] = abs32 ? 0.0 : b.origin[1];
      s.bounds[7] = abs32 ? 0.0 : b.origin[2];
    }
    auto* keys = gblw(static_cast<K*>(s.keys[dst]));  // the sort's first source buffer
    auto* vals = gblw(s.vals[dst]);
    auto* gend = gblw(s.gend);
    const auto* sx = gbl(s.x);
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const uint32_t i = tl.start + e;
      uint32_t l[3];
      for (int a = 0; a < 3; ++a)
        l[a] = quant_lane((sx[3 * (size_t)i + a] - b.bmin[a]) * b.scale, top);
      if (lanes_out) {
        lanes_out[3 * (size_t)i] = l[0];
        lanes_out[3 * (size_t)i + 1] = l[1];
        lanes_out[3 * (size_t)i + 2] = l[2];
      }
      keys[i] = T::code(l[0], l[1], l[2]);
      vals[i] = i;
      // a fresh build: group starts are marked by the tree build, which runs after the
      // sort (concurrently with the gather, which therefore leaves the marks alone)
      gend[i] = s.n == 1 ? 0u : kEnd;
      if (s.n == 1) s.parent_leaf[0] = kNoParent;
    }
  }
}

// Bounds only (used when keys are supplied by the caller).
__global__ void k_bounds(const SegDesc* __restrict__ segs, uint32_t nseg, int cell_bits) {
  const uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nseg) return;
  const SegDesc& s = segs[sidx];
  const Bounds b = seg_bounds(s.bbox, cell_bits);
  for (int a = 0; a < 3; ++a) {
    s.bounds[a] = b.bmin[a];
    s.bounds[5 + a] = s.precision == kPrecAbs32 ? 0.0 : b.origin[a];
  }
  s.bounds[3] = b.size;
  s.bounds[4] = b.scale;
}

__global__ void __launch_bounds__(kBlock)
k_iota(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
       int dst) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock)
      s.vals[dst][tl.start + e] = tl.start + e;
  }
}

// ----------------------------------------------------------------------------------
// 3. stable LSD radix sort, segmented reduce-then-scan (per pass: upsweep, scan,
//    downsweep) with B-bit digits (B = 8, 9 or 10: 2^B bins). Pure integer; reproduces
//    the (code, index) order of solvers/src/lbvh.rs:143-147 because the payload starts
//    as iota and every pass is stable, whatever the digit width.
// ----------------------------------------------------------------------------------

template <class K, int B>
__global__ void __launch_bounds__(kSortThreads)
k_radix_upsweep(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
                uint32_t* __restrict__ tile_hist, int shift, int src) {
  constexpr int BINS = 1 << B;
  __shared__ uint32_t hist[BINS];
  for (uint32_t t = xcd_block<kXcdBuildRun>(); t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const auto* keys = gbl(static_cast<const K*>(s.keys[src]) + tl.start);
    for (int d = threadIdx.x; d < BINS; d += kSortThreads) hist[d] = 0;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < tl.count; e += kSortThreads)
      atomicAdd(&hist[(uint32_t)(keys[e] >> shift) & (BINS - 1)], 1u);
    __syncthreads();
    for (int d = threadIdx.x; d < BINS; d += kSortThreads)
      tile_hist[(size_t)t * BINS + d] = hist[d];
    __syncthreads();
  }
}

// One block of 1024 threads per segment: thread (digit d, group g) scans one of the
// 1024 / 2^B groups of the segment's rows; row offsets = digit base (exclusive scan over
// digits of the digit totals) + exclusive prefix over earlier rows. The rows are the
// segment's tiles, or, for a segment of more than kScanChunkMin tiles, its chunks of
// kScanChunk tiles: their sums come from k_radix_chunk_sum and the offsets found here
// replace them in place (k_radix_chunk_offsets then spreads them over the tiles), so
// that no thread walks more than a few dozen rows in a row.
constexpr int kScanThreads = 1024;

template <int B>
__global__ void __launch_bounds__(kScanThreads)
k_radix_scan(const SegDesc* __restrict__ segs, uint32_t nseg,
             const uint32_t* __restrict__ tile_hist, uint32_t* __restrict__ tile_off,
             uint32_t* chunk_rows) {
  constexpr int BINS = 1 << B, CHUNKS = kScanThreads / BINS;
  __shared__ uint32_t part[CHUNKS][BINS];
  __shared__ uint32_t base[BINS];
  __shared__ uint32_t wsum[BINS / 64];
  const uint32_t d = threadIdx.x & (BINS - 1), g = threadIdx.x / BINS;
  for (uint32_t sidx = blockIdx.x; sidx < nseg; sidx += gridDim.x) {
    const SegDesc& s = segs[sidx];
    const bool chunked = s.scan_chunks > 0;
    const uint32_t* in = chunked ? chunk_rows : tile_hist;
    uint32_t* out = chunked ? chunk_rows : tile_off;  // chunk offsets replace their sums
    const uint32_t r0 = chunked ? s.scan_chunk0 : s.sort_tile0;
    const uint32_t nt = chunked ? s.scan_chunks : s.sort_tiles, per = (nt + CHUNKS - 1) / CHUNKS;
    const uint32_t c0 = r0 + min(nt, g * per), c1 = r0 + min(nt, (g + 1) * per);
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t t = c0; t < c1; ++t) sum += in[(size_t)t * BINS + d];
    part[g][d] = sum;
    __syncthreads();
    if (threadIdx.x < BINS) {
      uint32_t tot = 0;
      for (int q = 0; q < CHUNKS; ++q) tot += part[q][d];
      const int lane = d & 63;
      uint32_t inc = tot;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o);
        if (lane >= o) inc += v;
      }
      if (lane == 63) wsum[d >> 6] = inc;
      base[d] = inc - tot;
    }
    __syncthreads();
    uint32_t run = base[d];
    for (uint32_t w = 0; w < (d >> 6); ++w) run += wsum[w];
    for (uint32_t q = 0; q < g; ++q) run += part[q][d];
#pragma unroll 8
    for (uint32_t t = c0; t < c1; ++t) {  // each row read before this thread overwrites it
      const uint32_t h = in[(size_t)t * BINS + d];
      out[(size_t)t * BINS + d] = run;
      run += h;
    }
    __syncthreads();
  }
}

// The chunks of the segments scanned in chunks: Tile{segment, first sort tile (of the
// batch), tiles, -}, chunk c's row at c * 2^B.
constexpr int kChunkThreads = 256;

template <int B>
__global__ void __launch_bounds__(kChunkThreads)
k_radix_chunk_sum(const Tile* __restrict__ chunks, uint32_t nchunks,
                  const uint32_t* __restrict__ tile_hist, uint32_t* __restrict__ chunk_rows) {
  constexpr int BINS = 1 << B;
  for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const Tile ch = chunks[c];
    for (uint32_t d = threadIdx.x; d < (uint32_t)BINS; d += kChunkThreads) {
      uint32_t sum = 0;
#pragma unroll 8
      for (uint32_t t = ch.start; t < ch.start + ch.count; ++t) sum += tile_hist[(size_t)t * BINS + d];
      chunk_rows[(size_t)c * BINS + d] = sum;
    }
  }
}

template <int B>
__global__ void __launch_bounds__(kChunkThreads)
k_radix_chunk_offsets(const Tile* __restrict__ chunks, uint32_t nchunks,
                      const uint32_t* __restrict__ tile_hist,
                      const uint32_t* __restrict__ chunk_rows, uint32_t* __restrict__ tile_off) {
  constexpr int BINS = 1 << B;
  for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const Tile ch = chunks[c];
    for (uint32_t d = threadIdx.x; d < (uint32_t)BINS; d += kChunkThreads) {
      uint32_t run = chunk_rows[(size_t)c * BINS + d];
#pragma unroll 8
      for (uint32_t t = ch.start; t < ch.start + ch.count; ++t) {
        const uint32_t h = tile_hist[(size_t)t * BINS + d];
        tile_off[(size_t)t * BINS + d] = run;
        run += h;
      }
    }
  }
}

template <class K, int B>
__global__ void __launch_bounds__(kSortThreads)
k_radix_downsweep(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
                  const uint32_t* __restrict__ tile_off, int shift, int src) {
  constexpr int BINS = 1 << B;
  constexpr int W = kSortThreads / 64;
  constexpr int PER_WAVE = kSortTile / W;
  constexpr int R = (BINS + kSortThreads - 1) / kSortThreads;  // digits per thread, tile scan
  __shared__ uint32_t wcnt[W][BINS];
  __shared__ uint32_t dstart[BINS];
  __shared__ K skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  __shared__ uint32_t shw[W];  // block_exclusive_scan's per-wave totals
  static_assert(kSortThreads % 64 == 0 && W * 64 == kSortThreads,
                "one shw entry per wave of the block");
  static_assert(R * kSortThreads >= BINS, "the tile scan covers every digit");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (uint32_t t = xcd_block<kXcdBuildRun>(); t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const auto* keys_in = gbl(static_cast<const K*>(s.keys[src]) + tl.start);
    const auto* vals_in = gbl(s.vals[src] + tl.start);
    auto* keys_out = gblw(static_cast<K*>(s.keys[src ^ 1]));
    auto* vals_out = gblw(s.vals[src ^ 1]);
    for (int k = 0; k < W; ++k)
      for (int d = threadIdx.x; d < BINS; d += kSortThreads) wcnt[k][d] = 0;
    __syncthreads();

    K key[kSortItems];
    uint32_t val[kSortItems], dig[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      const uint32_t e = w * PER_WAVE + j * 64 + lane;  // striped: index order = (j, lane)
      const bool valid = e < tl.count;
      key[j] = valid ? keys_in[e] : (K)0;
      val[j] = valid ? vals_in[e] : 0u;
      const uint32_t dd = valid ? ((uint32_t)(key[j] >> shift) & (BINS - 1)) : (uint32_t)BINS;
      // lanes holding the same digit (bit B separates the invalid tail)
      unsigned long long peers = ~0ull;
#pragma unroll
      for (int b = 0; b <= B; ++b) {
        const bool bit = (dd >> b) & 1u;
        const unsigned long long bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      uint32_t r = 0;
      if (valid) {
        const uint32_t before = wcnt[w][dd];
        r = before + (uint32_t)__popcll(peers & lt);
        if ((peers & lt) == 0ull) wcnt[w][dd] = before + (uint32_t)__popcll(peers);
      }
      dig[j] = dd;
      rank[j] = r;
    }
    __syncthreads();
    {
      // thread q owns digits [q R, q R + R): per-wave starts, then the tile's digit starts
      uint32_t tot[R], sum = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int d = threadIdx.x * R + r;
        uint32_t run = 0;
        for (int k = 0; k < W && d < BINS; ++k) {
          const uint32_t c = wcnt[k][d];
          wcnt[k][d] = run;
          run += c;
        }
        tot[r] = run;
        sum += run;
      }
      uint32_t first = block_exclusive_scan(sum, shw);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (threadIdx.x * R + r < BINS) dstart[threadIdx.x * R + r] = first;
        first += tot[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      if (dig[j] < (uint32_t)BINS) {
        const uint32_t p = dstart[dig[j]] + wcnt[w][dig[j]] + rank[j];
        skey[p] = key[j];
        sval[p] = val[j];
      }
    }
    __syncthreads();
    const uint32_t* off = tile_off + (size_t)t * BINS;
    for (uint32_t p = threadIdx.x; p < tl.count; p += kSortThreads) {
      const K k = skey[p];
      const uint32_t d = (uint32_t)(k >> shift) & (BINS - 1);
      const uint32_t dst = off[d] + (p - dstart[d]);
      keys_out[dst] = k;
      vals_out[dst] = sval[p];
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------------
// 4. gather sorted leaves: origin-relative f32 position + mass
//    (solvers/src/lbvh.rs:453-456; gpu/src/fused_core.rs:32-45)
// ----------------------------------------------------------------------------------

// Sorted leaves of a fresh build, or the positions of a cached tree refreshed in its
// existing leaf order (the same arithmetic). 1024 threads per 2048-element tile keep more
// of the random reads in flight.
constexpr int kGatherThreads = 1024;

__global__ void __launch_bounds__(kGatherThreads)
k_gather(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  for (uint32_t t = xcd_block<kXcdBuildRun>(); t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const double o0 = s.bounds[5], o1 = s.bounds[6], o2 = s.bounds[7];
    for (uint32_t e = threadIdx.x; e < tl.count; e += kGatherThreads) {
      const uint32_t k = tl.start + e;
      const size_t i = min(gbl(s.vals[0])[k], s.n - 1);
      const auto* x = gbl(s.x);
      auto* lp = gblw(reinterpret_cast<float*>(s.leaf)) + 4 * (size_t)k;
      lp[0] = (float)(x[3 * i] - o0);
      lp[1] = (float)(x[3 * i + 1] - o1);
      lp[2] = (float)(x[3 * i + 2] - o2);
      lp[3] = (float)gbl(s.m)[i];
    }
  }
}

// ----------------------------------------------------------------------------------
// 5. Karras radix-tree build (solvers/src/lbvh.rs:153-240,276-306), signed int
//    indices; each internal node also records its leaf range and publishes itself
//    as the right child of its split position, from which the skip pointers follow.
// ----------------------------------------------------------------------------------

// The sorted keys of a tile and kKarrasPad on either side, staged in LDS: the searches
// of most nodes stay inside this window and only the long ranges of nodes near the root
// read global memory.
constexpr int kKarrasPad = 512;

template <class K>
struct KeyWindow {
  const K* codes;
  const K* win;
  int w0, wlen;
  // The window as an LDS read (clamped index) and the global key only outside it, so
  // neither becomes a flat load through a selected pointer.
  __device__ __forceinline__ K at(int g) const {
    const int q = g - w0;
    const bool in = q >= 0 && q < wlen;
    K v = win[in ? q : 0];
    if (!in) v = gbl(codes)[g];
    return v;
  }
};

// delta(a, b) of the node's own key ka (position a) and the key at b; -1 outside [0, n)
template <class K>
__device__ __forceinline__ int kdelta(const KeyWindow<K>& kw, int n, int a, K ka, int b) {
  if (b < 0 || b >= n) return -1;
  const K cb = kw.at(b);
  if (ka == cb) return KeyTraits<K>::tie_base + __clz((int)((uint32_t)a ^ (uint32_t)b));
  return KeyTraits<K>::clz(ka ^ cb);
}

// 1024 threads per 2048-key tile: two nodes per thread, so the binary searches of a tile
// are spread over 16 waves that hide each other's LDS and memory latency.
constexpr int kKarrasThreads = 1024;

template <class K>
__global__ void __launch_bounds__(kKarrasThreads)
k_karras(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  __shared__ K win[kSortTile + 2 * kKarrasPad];
  for (uint32_t t = xcd_block<kXcdBuildRun>(); t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const int n = (int)s.n;
    const K* codes = static_cast<const K*>(s.keys[0]);
    KeyWindow<K> kw{codes, win, (int)tl.start - kKarrasPad, (int)tl.count + 2 * kKarrasPad};
    for (int q = threadIdx.x; q < kw.wlen; q += kKarrasThreads) {
      const int g = kw.w0 + q;
      if (g >= 0 && g < n) win[q] = codes[g];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < tl.count; e += kKarrasThreads) {
      const int i = (int)(tl.start + e);
      if (i >= n - 1) continue;
      const K ki = kw.at(i);
      const int dir = kdelta(kw, n, i, ki, i + 1) > kdelta(kw, n, i, ki, i - 1) ? 1 : -1;
      const int dmin = kdelta(kw, n, i, ki, i - dir);
      int lmax = 2;
      while (kdelta(kw, n, i, ki, i + lmax * dir) > dmin) lmax *= 2;
      int l = 0;
      for (int tt = lmax / 2; tt > 0; tt /= 2)
        if (kdelta(kw, n, i, ki, i + (l + tt) * dir) > dmin) l += tt;
      const int j = i + l * dir;
      const int dnode = kdelta(kw, n, i, ki, j);
      int sp = 0, tt = l;
      for (;;) {
        tt = (tt + 1) / 2;
        if (kdelta(kw, n, i, ki, i + (sp + tt) * dir) > dnode) sp += tt;
        if (tt <= 1) break;
      }
      const int gamma = i + sp * dir + (dir < 0 ? dir : 0);
      const int lo = i < j ? i : j, hi = i < j ? j : i;
      const uint32_t lu = (lo == gamma) ? (uint32_t)gamma : (uint32_t)(n + gamma);
      const uint32_t ru = (hi == gamma + 1) ? (uint32_t)(gamma + 1) : (uint32_t)(n + gamma + 1);
      const uint32_t me = (uint32_t)(n + i);
      // node i's own fields; its parent field is written by its parent (the root's here)
      uint32_t* tp = reinterpret_cast<uint32_t*>(s.topo);
      tp[4 * (size_t)i] = (uint32_t)lo;
      tp[4 * (size_t)i + 1] = (uint32_t)hi;
      tp[4 * (size_t)i + 2] = enc_children(lu, ru, (uint32_t)n);
      if (lu < (uint32_t)n) s.parent_leaf[lu] = me; else tp[4 * (size_t)(lu - n) + 3] = me;
      if (ru < (uint32_t)n) s.parent_leaf[ru] = me; else tp[4 * (size_t)(ru - n) + 3] = me;
      s.rs[gamma] = ru;
      if (i == 0) tp[3] = kNoParent;
      // walk groups: the subtrees of at most group_size (<= 64) bodies whose parent holds
      // more; a node above the limit marks those of its two children
      const int gmax = (int)s.group_size;
      if (hi - lo + 1 > gmax) {
        if (gamma - lo + 1 <= gmax) s.gend[lo] = (uint32_t)gamma;
        if (hi - gamma <= gmax) s.gend[gamma + 1] = (uint32_t)hi;
      } else if (i == 0) {
        s.gend[0] = (uint32_t)(n - 1);
      }
    }
    __syncthreads();  // the window is refilled for the next tile
  }
}

// ----------------------------------------------------------------------------------
// 6. bottom-up aggregate (fold order fixed: left, right; solvers/src/lbvh.rs:623-636),
//    in passes that need only workgroup-scope ordering:
//    (a) per tile of kAggTile sorted leaves, one block climbs from every leaf with
//        arrival counters in LDS and keeps in-tile node aggregates in LDS; a climb
//        stops below the first node whose leaf range leaves the tile and hands that
//        subtree root on, in the tile's own list;
//    (b) for systems of more than kMidSpan bodies, one block per super-tile of
//        kMidTiles tiles climbs from its tiles' roots through the nodes inside the
//        super-tile, with arrival counters in global memory, and hands on its own roots;
//    (c) one block per segment climbs from the handed-over roots to the root.
//    Each completed node is written once as its walk record (and quadrupole). A node's
//    record depends only on its children's, so the records are the same whichever pass
//    completes a node and in whatever order its children arrive.
// ----------------------------------------------------------------------------------

// 16-byte record i of a float4 array as one global load (a flat load would also hold up
// the LDS waits that follow it).
__device__ __forceinline__ float4 gload4(const float4* p, size_t i) {
  const auto* q = gbl(reinterpret_cast<const float*>(p)) + 4 * i;
  return make_float4(q[0], q[1], q[2], q[3]);
}

// 16 B at a 32-bit byte offset from a wave-uniform base: the global load's scalar-base
// form (base in SGPRs, offset in one VGPR), with no 64-bit address arithmetic per lane
__device__ __forceinline__ float4 gload4_at(const char* base, uint32_t off) {
  const auto* q = gbl(reinterpret_cast<const float*>(base + off));
  return make_float4(q[0], q[1], q[2], q[3]);
}

__device__ __forceinline__ uint4 gload_u4(const uint4* p, size_t i) {
  const auto* q = gbl(reinterpret_cast<const uint32_t*>(p)) + 4 * i;
  return make_uint4(q[0], q[1], q[2], q[3]);
}

struct AggV {
  float3 com, mn, mx;
  float mass;
  float q[6];  // traceless quadrupole xx, yy, zz, xy, xz, yz about com
};

__device__ __forceinline__ AggV leaf_agg(float4 p) {
  AggV r;
  r.com = make_float3(p.x, p.y, p.z);
  r.mn = r.com;
  r.mx = r.com;
  r.mass = p.w;
  for (int c = 0; c < 6; ++c) r.q[c] = 0.f;
  return r;
}

// The fold and the record below are written out without floating-point contraction, so
// a node's record is the same bits whichever kernel computes it (the compiler would
// otherwise fuse multiplies and adds differently in each inlining context).
template <bool QUAD>
__device__ __forceinline__ AggV fold_agg(const AggV& l, const AggV& r) {
#pragma clang fp contract(off)
  AggV o;
  o.mass = l.mass + r.mass;
  if (o.mass > 0.0f) {
    o.com.x = (l.com.x * l.mass + r.com.x * r.mass) / o.mass;
    o.com.y = (l.com.y * l.mass + r.com.y * r.mass) / o.mass;
    o.com.z = (l.com.z * l.mass + r.com.z * r.mass) / o.mass;
  } else {
    o.com.x = (l.mn.x + l.mx.x + r.mn.x + r.mx.x) * 0.25f;
    o.com.y = (l.mn.y + l.mx.y + r.mn.y + r.mx.y) * 0.25f;
    o.com.z = (l.mn.z + l.mx.z + r.mn.z + r.mx.z) * 0.25f;
  }
  o.mn = make_float3(fminf(l.mn.x, r.mn.x), fminf(l.mn.y, r.mn.y), fminf(l.mn.z, r.mn.z));
  o.mx = make_float3(fmaxf(l.mx.x, r.mx.x), fmaxf(l.mx.y, r.mx.y), fmaxf(l.mx.z, r.mx.z));
  for (int c = 0; c < 6; ++c) o.q[c] = 0.f;
  if (QUAD && o.mass > 0.0f) {
    // parallel-axis shift of each child's traceless quadrupole to the parent com
    const AggV* ch[2] = {&l, &r};
    for (int k = 0; k < 2; ++k) {
      const AggV& c = *ch[k];
      const float dx = c.com.x - o.com.x, dy = c.com.y - o.com.y, dz = c.com.z - o.com.z;
      const float d2 = dx * dx + dy * dy + dz * dz, m = c.mass;
      o.q[0] += c.q[0] + m * (3.f * dx * dx - d2);
      o.q[1] += c.q[1] + m * (3.f * dy * dy - d2);
      o.q[2] += c.q[2] + m * (3.f * dz * dz - d2);
      o.q[3] += c.q[3] + m * 3.f * dx * dy;
      o.q[4] += c.q[4] + m * 3.f * dx * dz;
      o.q[5] += c.q[5] + m * 3.f * dy * dz;
    }
  }
  return o;
}

// Walk record of internal node j from its aggregate (solvers/src/lbvh.rs:602-611 for
// centre/half/delta; the acceptance radius follows from :511-514:
// theta*(d - delta) >= s  <=>  d >= s/theta + delta).
// The walk record of internal node j (topology fields given: leaf range [lo, hi],
// children, and the skip pointer nx = rs[hi] or kEnd).
template <bool QUAD>
__device__ __forceinline__ void store_node(const SegDesc& s, uint32_t j, const AggV& r,
                                           float inv_theta, uint32_t lo, uint32_t hi,
                                           uint32_t cl, uint32_t cr, uint32_t nx,
                                           bool write_box = true) {
#pragma clang fp contract(off)
  const float cx = (r.mn.x + r.mx.x) * 0.5f, cy = (r.mn.y + r.mx.y) * 0.5f,
              cz = (r.mn.z + r.mx.z) * 0.5f;
  const float hx = (r.mx.x - r.mn.x) * 0.5f, hy = (r.mx.y - r.mn.y) * 0.5f,
              hz = (r.mx.z - r.mn.z) * 0.5f;
  const float size = 2.0f * fmaxf(hx, fmaxf(hy, hz));
  const float ddx = r.com.x - cx, ddy = r.com.y - cy, ddz = r.com.z - cz;
  const float delta = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
  const float rcrit = size * inv_theta + delta;
  Node nd;
  nd.a = make_float4(r.com.x, r.com.y, r.com.z, r.mass);
  const float ch = __uint_as_float(enc_entries(cl, cr, s.n));
  NodeBox bx;
  bx.lo = make_float4(r.mn.x, r.mn.y, r.mn.z, __uint_as_float(lo));
  bx.hi = make_float4(r.mx.x, r.mx.y, r.mx.z, __uint_as_float(hi));
  if (QUAD) {
    // the group walk's classifier radii: squared monopole-band radius (s / (theta
    // mono_ratio) + delta; FLT_MAX when the band is off) and the node's radius about its
    // com (far-field test); the skip pointer is rebuilt from rs[hi] where needed
    const float ex = fmaxf(r.com.x - r.mn.x, r.mx.x - r.com.x),
                ey = fmaxf(r.com.y - r.mn.y, r.mx.y - r.com.y),
                ez = fmaxf(r.com.z - r.mn.z, r.mx.z - r.com.z);
    const float rn = sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
    const float rm = fmaf(size, s.mono_inv, delta);
    nd.h = make_float4(rcrit * rcrit, ch, s.mono_inv > 0.f ? rm * rm : 3.4e38f, rn);
  } else {
    nd.h = make_float4(rcrit * rcrit, ch, __uint_as_float(nx), 0.f);
  }
  s.node[j] = nd;
  if (write_box) s.box[j] = bx;
  if (QUAD) {
    s.quad[2 * j] = make_float4(r.q[0], r.q[1], r.q[2], r.q[3]);
    s.quad[2 * j + 1] = make_float4(r.q[4], r.q[5], 0.f, 0.f);
  }
}

template <bool QUAD>
__device__ __forceinline__ AggV global_child(const SegDesc& s, uint32_t c) {
  const uint32_t n = s.n;
  if (c < n) return leaf_agg(s.leaf[c]);
  AggV r;
  const float4 a = s.node[c - n].a, b = s.box[c - n].lo, cc = s.box[c - n].hi;
  r.com = make_float3(a.x, a.y, a.z);
  r.mass = a.w;
  r.mn = make_float3(b.x, b.y, b.z);
  r.mx = make_float3(cc.x, cc.y, cc.z);
  if (QUAD) {
    const float4 q0 = s.quad[2 * (c - n)], q1 = s.quad[2 * (c - n) + 1];
    r.q[0] = q0.x; r.q[1] = q0.y; r.q[2] = q0.z; r.q[3] = q0.w; r.q[4] = q1.x; r.q[5] = q1.y;
  } else {
    for (int k = 0; k < 6; ++k) r.q[k] = 0.f;
  }
  return r;
}

// Tile pass, one block of one thread per leaf per tile, in two phases:
//  A. every leaf climbs with arrival words in LDS, as far as the nodes inside the tile
//     go; an arrival only adds its subtree height, and the second arrival at a node
//     learns the node's height (1 + the taller child's) and climbs on. The nodes that
//     stop a climb by leaving the tile get the climb's subtree root handed on.
//  B. the in-tile nodes, bucketed by height in LDS, are folded one height at a time
//     (every child of a height-h node is lower), so the lanes that fold are packed into
//     as few waves as the bucket needs: the fold, most of a climb step's instructions,
//     runs about once per node instead of once per wave per tree level.
// Topology is staged in LDS first by coalesced loads (an in-tile node j has t0 <= j < t1:
// an internal node's leaf range has its own index at one end), packed tile-relative in
// 8 bytes; each leaf is staged into its parent's LDS record, which the parent's fold
// reads before overwriting it with its own aggregate.
constexpr int kAggThreads = kAggTile;
// Heights of in-tile nodes stay below this: the common prefix of a node's range (key bits,
// then index bits for equal keys: at most 64 + 32) grows strictly from parent to child.
constexpr uint32_t kMaxHeight = 128;
// packed in-tile topology (s_info): x = lo | hi << 10 | inside | parent << 21 | parent in
// tile; y = split | root | leaf flags (kLeftLeaf, kRightLeaf), all tile-relative
constexpr uint32_t kInInside = 1u << 20, kInParent = 1u << 31, kInRoot = 1u << 20;

// Block barrier that orders LDS only: the global stores before it (node records, read by
// later kernels only) are not waited for, as __syncthreads' release fence would.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool QUAD>
__device__ __forceinline__ AggV lds_child(const float4 (*agg)[kAggTile], uint32_t c) {
  AggV r;
  const float4 p0 = agg[0][c], p1 = agg[1][c], p2 = agg[2][c];
  r.com = make_float3(p0.x, p0.y, p0.z);
  r.mass = p0.w;
  r.mn = make_float3(p1.x, p1.y, p1.z);
  r.mx = make_float3(p2.x, p2.y, p2.z);
  if (QUAD) {
    const float4 p3 = agg[QUAD ? 3 : 0][c];
    r.q[0] = p1.w; r.q[1] = p2.w; r.q[2] = p3.x; r.q[3] = p3.y; r.q[4] = p3.z; r.q[5] = p3.w;
  } else {
    for (int k = 0; k < 6; ++k) r.q[k] = 0.f;
  }
  return r;
}

// Two blocks per CU (the LDS above fits twice into 160 KB): 8 waves per SIMD, so at most
// 64 VGPRs.
template <bool QUAD>
__global__ void __launch_bounds__(kAggThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_aggregate_tile(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles,
                 uint32_t ntiles, int group) {
  constexpr int NP = QUAD ? 4 : 3;  // com+mass, min(+q0), max(+q1), q2..q5
  __shared__ float4 s_agg[NP][kAggTile];
  __shared__ uint2 s_info[kAggTile];
  __shared__ uint32_t s_hw[kAggTile];  // A: arrivals << 16 | height; B: nodes by height
  __shared__ uint32_t s_hoff[kMaxHeight + 1], s_hcur[kMaxHeight];
  __shared__ uint32_t s_nroots, s_maxh;
  const uint32_t e = threadIdx.x;  // this thread's leaf (and staged node) in the tile
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const uint32_t n = s.n;
    if (n < 2) continue;
    const uint32_t t0 = tl.start, cnt = tl.count, t1 = t0 + cnt;
    const float inv_theta = 1.0f / s.theta;
    // every node's box record is read by the per-target walk and by the group walk above
    // theta = 1 (compared as bits, as the walk does) or with buckets
    const bool box_all = !group || __float_as_uint(s.theta) > __float_as_uint(1.0f) ||
                         s.leaf_size > 1u;
    if (e < kMaxHeight) s_hcur[e] = 0u;
    if (e == 0) s_nroots = 0u;
    float4 lf = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t p = kNoParent;
    if (e < cnt) {
      lf = gload4(s.leaf, t0 + e);
      p = gbl(s.parent_leaf)[t0 + e];
      uint2 inf = make_uint2(0u, 0u);
      if (t0 + e < n - 1) {
        const uint4 tj = gload_u4(s.topo, t0 + e);
        if (tj.x >= t0 && tj.y < t1) {
          const uint32_t pr = tj.w - n - t0;  // wraps far past cnt for kNoParent
          inf.x = (tj.x - t0) | (tj.y - t0) << 10 | kInInside;
          if (tj.w != kNoParent && pr < cnt) inf.x |= pr << 21 | kInParent;
          inf.y = ((tj.z & kSplitMask) - t0) | (tj.z & (kLeftLeaf | kRightLeaf)) |
                  (tj.w == kNoParent ? kInRoot : 0u);
        }
      }
      s_info[e] = inf;
      s_hw[e] = 0u;
    }
    lds_barrier();
    // A: climbs (each in-tile node is reached twice; the bound only guards corrupted input)
    if (e < cnt) {
      uint32_t u = t0 + e, hu = 0;
      for (uint32_t guard = 0; guard < cnt; ++guard) {
        const uint32_t q = p - n - t0;
        const uint32_t ix = q < cnt ? s_info[q].x : 0u;
        if (!(ix & kInInside)) {  // u is a maximal subtree of the tile
          s.roots[t0 + atomicAdd(&s_nroots, 1u)] = u;
          break;
        }
        if (hu == 0u)  // u is a leaf: stage it for its parent's fold (left child = split)
          s_agg[(s_info[q].y & 1023u) == e ? 0 : 1][q] = lf;
        const uint32_t old = atomicAdd(&s_hw[q], 0x10000u + hu + 1u);
        if (old < 0x10000u) break;  // first arrival
        const uint32_t h = max(old & 0xFFFFu, hu + 1u);
        const uint32_t hb = h < kMaxHeight ? h : kMaxHeight - 1;
        s_hw[q] = hb;
        atomicAdd(&s_hcur[hb], 1u);
        hu = h;
        u = n + t0 + q;
        if (ix & kInParent) {
          p = n + t0 + ((ix >> 21) & 1023u);
          continue;
        }
        if (!(s_info[q].y & kInRoot)) s.roots[t0 + atomicAdd(&s_nroots, 1u)] = u;
        break;
      }
    }
    lds_barrier();
    // the in-tile nodes in buckets by height: offsets, then each node's slot
    const uint32_t hn = (e < cnt && (s_info[e].x & kInInside)) ? s_hw[e] : 0u;
    if (e < 64) {
      const uint32_t c0 = s_hcur[2 * e], c1 = s_hcur[2 * e + 1], v = c0 + c1;
      uint32_t inc = v;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t w = __shfl_up(inc, o);
        if ((int)e >= o) inc += w;
      }
      s_hoff[2 * e] = inc - v;
      s_hoff[2 * e + 1] = inc - v + c0;
      s_hcur[2 * e] = inc - v;
      s_hcur[2 * e + 1] = inc - v + c0;
      if (e == 63) s_hoff[kMaxHeight] = inc;
      // the tallest height present: the last non-empty bucket
      const unsigned long long any = __ballot(v != 0u);
      if (e == 0) s_maxh = any ? 2u * (63u - (uint32_t)__clzll(any)) + 1u : 0u;
    }
    lds_barrier();
    if (hn) s_hw[atomicAdd(&s_hcur[hn], 1u)] = e;
    lds_barrier();
    // B: folds, one height at a time
    const uint32_t H = s_maxh;
    for (uint32_t h = 1; h <= H; ++h) {
      const uint32_t idx = s_hoff[h] + e;
      if (idx < s_hoff[h + 1]) {
        const uint32_t q = s_hw[idx];
        const uint2 inf = s_info[q];
        const uint32_t g = inf.y & 1023u;
        const bool ll = inf.y & kLeftLeaf, rl = inf.y & kRightLeaf;
        const AggV a = ll ? leaf_agg(s_agg[0][q]) : lds_child<QUAD>(s_agg, g);
        const AggV b = rl ? leaf_agg(s_agg[1][q]) : lds_child<QUAD>(s_agg, g + 1);
        const AggV r = fold_agg<QUAD>(a, b);
        s_agg[0][q] = make_float4(r.com.x, r.com.y, r.com.z, r.mass);
        s_agg[1][q] = make_float4(r.mn.x, r.mn.y, r.mn.z, r.q[0]);
        s_agg[2][q] = make_float4(r.mx.x, r.mx.y, r.mx.z, r.q[1]);
        if (QUAD) s_agg[QUAD ? 3 : 0][q] = make_float4(r.q[2], r.q[3], r.q[4], r.q[5]);
        const uint32_t lo = t0 + (inf.x & 1023u), hi = t0 + ((inf.x >> 10) & 1023u);
        const uint32_t cl = ll ? t0 + g : n + t0 + g, cr = rl ? t0 + g + 1 : n + t0 + g + 1;
        const uint32_t nx = QUAD ? 0u : (hi == n - 1 ? kEnd : gbl(s.rs)[hi]);
        // box records: every node's for the builds that read them (per-target walk, the
        // group walk above theta = 1 or with buckets); otherwise only those of the roots
        // handed to the climbing passes (no in-tile parent), which fold their boxes
        const bool in_parent = (inf.x & kInParent) &&
                               (s_info[(inf.x >> 21) & 1023u].x & kInInside);
        store_node<QUAD>(s, t0 + q, r, inv_theta, lo, hi, cl, cr, nx, box_all || !in_parent);
      }
      lds_barrier();
    }
    if (e == 0) s.rcount[t0 / kAggTile] = s_nroots;
    lds_barrier();
  }
}

constexpr int kTopThreads = 1024;

// The handed-over roots of one pass, as lists: list i holds count[i] roots at
// roots[i * span ...]. `pre` (in LDS) receives the exclusive prefix of the counts; the
// block's lanes then take roots one at a time from a workgroup counter over the
// concatenation, so the work is balanced whatever the list lengths.
__device__ uint32_t root_prefix(const uint32_t* __restrict__ count, uint32_t nlists,
                                uint32_t* pre, uint32_t* sh_wave, uint32_t* sh_carry) {
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nlists; c0 += blockDim.x) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < nlists ? count[i] : 0u;
    const uint32_t ex = block_exclusive_scan(v, sh_wave);
    if (i < nlists) pre[i] = carry + ex;
    if (threadIdx.x == blockDim.x - 1) *sh_carry = carry + ex + v;
    __syncthreads();
    carry = *sh_carry;
    __syncthreads();
  }
  return carry;
}

// The list holding concatenated root r: the last i with pre[i] <= r.
__device__ __forceinline__ uint32_t root_list(const uint32_t* pre, uint32_t nlists, uint32_t r) {
  uint32_t lo = 0, hi = nlists - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One climb step of the global passes at node j = p - n: the arrival counter in global
// memory at workgroup scope (only this block reaches the node), the second arrival folds
// the children's records and writes the node's. Returns the parent to go on with, or
// kNoParent when this lane arrived first.
template <bool QUAD>
__device__ __forceinline__ uint32_t global_climb(const SegDesc& s, uint32_t j, const uint4& tj,
                                                 float inv_theta) {
  const uint32_t n = s.n;
  const uint32_t lo = tj.x, hi = tj.y, cl = left_child(tj.z, n), cr = right_child(tj.z, n);
  const uint32_t old = __hip_atomic_fetch_add(&s.flag[j], 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
  if (old == 0u) return kNoParent;
  __hip_atomic_store(&s.flag[j], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const uint32_t nx = hi == n - 1 ? kEnd : gbl(s.rs)[hi];
  const AggV a = fold_agg<QUAD>(global_child<QUAD>(s, cl), global_child<QUAD>(s, cr));
  store_node<QUAD>(s, j, a, inv_theta, lo, hi, cl, cr, nx);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  return tj.w;
}

// Middle pass, for systems of more than kMidSpan bodies: one block per super-tile of
// kMidSpan sorted leaves (kMidTiles tiles) climbs from the roots its tiles handed over,
// through the nodes whose leaf range stays inside the super-tile (arrival counters in
// global memory at workgroup scope: no other block reaches them), and hands the roots of
// its maximal subtrees to the top pass. The top pass is then left with the nodes that
// straddle super-tiles instead of all those that straddle tiles: on a 2^23-body system
// one block per super-tile runs instead of one block for the whole cross-tile tree.
// Both climbing passes are held to 64 VGPRs (8 waves per SIMD), so that two 1024-thread
// blocks share a CU: at 69 VGPRs one block per CU made the ensemble's top pass 35% slower.
template <bool QUAD>
__global__ void __launch_bounds__(kTopThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_aggregate_mid(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles,
                uint32_t ntiles) {
  __shared__ uint32_t pre[kMidTiles];
  __shared__ uint32_t sh_wave[kTopThreads / 64], sh_carry;
  __shared__ uint32_t next_root, s_nout;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const uint32_t n = s.n;
    const uint32_t S0 = tl.start, S1 = tl.start + tl.count;
    const uint32_t f = S0 / kAggTile, nl = (tl.count + kAggTile - 1) / kAggTile;
    if (threadIdx.x == 0) {
      next_root = 0u;
      s_nout = 0u;
    }
    const uint32_t total = root_prefix(s.rcount + f, nl, pre, sh_wave, &sh_carry);
    const float inv_theta = 1.0f / s.theta;
    uint32_t p = kNoParent, cur = 0;
    for (uint32_t guard = 0; guard < total + tl.count; ++guard) {
      if (p == kNoParent) {
        const uint32_t r = atomicAdd(&next_root, 1u);
        if (r >= total) break;
        const uint32_t i = root_list(pre, nl, r);
        cur = s.roots[(f + i) * kAggTile + (r - pre[i])];
        p = cur < n ? s.parent_leaf[cur] : s.topo[cur - n].w;
        continue;
      }
      const uint32_t j = p - n;
      if (j >= n - 1) {
        p = kNoParent;
        continue;
      }
      const uint4 tj = gload_u4(s.topo, j);
      if (tj.x < S0 || tj.y >= S1) {  // leaves the super-tile: cur goes to the top pass
        const uint32_t slot = atomicAdd(&s_nout, 1u);
        s.roots2[S0 + slot] = cur;
        p = kNoParent;
        continue;
      }
      const uint32_t up = global_climb<QUAD>(s, j, tj, inv_theta);
      if (up != kNoParent) cur = p;
      p = up;
    }
    __syncthreads();
    if (threadIdx.x == 0) s.rcount2[S0 / kMidSpan] = s_nout;
    __syncthreads();
  }
}

// Top pass: one block per system climbs from the roots handed over by the tile pass (or,
// for systems of more than kMidSpan bodies, by the middle pass) to the root. Each lane
// climbs from one root at a time and takes the next from a workgroup counter as soon as
// its climb ends (first arrival at a node), so the block stays busy and the critical path
// is the height of the part of the tree above the tiles (super-tiles).
constexpr uint32_t kTopLists = ((uint32_t)1 << 29) / kMidSpan;  // lists of the largest system

template <bool QUAD>
__global__ void __launch_bounds__(kTopThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_aggregate_top(const SegDesc* __restrict__ segs, uint32_t nseg) {
  __shared__ uint32_t pre[kTopLists];
  __shared__ uint32_t sh_wave[kTopThreads / 64], sh_carry;
  __shared__ uint32_t next_root;
  for (uint32_t sidx = blockIdx.x; sidx < nseg; sidx += gridDim.x) {
    const SegDesc& s = segs[sidx];
    const uint32_t n = s.n;
    if (threadIdx.x == 0) next_root = 0u;
    __syncthreads();
    if (n >= 2) {
      const bool mid = n > kMidMin;
      const uint32_t span = mid ? kMidSpan : (uint32_t)kAggTile;
      const uint32_t nl = (n + span - 1) / span;  // <= kMidMin / kAggTile without the middle pass
      const uint32_t* roots = mid ? s.roots2 : s.roots;
      const uint32_t total = root_prefix(mid ? s.rcount2 : s.rcount, nl, pre, sh_wave, &sh_carry);
      const float inv_theta = 1.0f / s.theta;
      uint32_t p = kNoParent;
      // a lane takes at most `total` roots and climbs at most n-1 nodes in all
      for (uint32_t guard = 0; guard < total + n; ++guard) {
        if (p == kNoParent) {
          const uint32_t r = atomicAdd(&next_root, 1u);
          if (r >= total) break;
          const uint32_t i = root_list(pre, nl, r);
          const uint32_t u = roots[i * span + (r - pre[i])];
          p = u < n ? s.parent_leaf[u] : s.topo[u - n].w;
          continue;
        }
        const uint32_t j = p - n;
        if (j >= n - 1) {
          p = kNoParent;
          continue;
        }
        // topology comes from the tree build (an earlier kernel): load it with the arrival
        p = global_climb<QUAD>(s, j, gload_u4(s.topo, j), inv_theta);
      }
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------------
// 6b. flatten, exported for parity only (the walks do not need it): the reference's DFS
//     pre-order layout (solvers/src/lbvh.rs:545-614), every node placed in parallel. In
//     pre-order with the left child first, the nodes before u are its ancestors and the
//     subtrees hanging left of its root path; those cover the leaves [0, lo) and hold
//     2 lo minus their count of nodes, one per right edge of the path, so
//     slot(u) = 2 lo + (left edges on the root path of u).
// ----------------------------------------------------------------------------------

__device__ uint32_t flat_slot(const SegDesc& s, uint32_t u) {
  const uint32_t n = s.n;
  if (n == 1) return 0u;
  const uint32_t lo = u < n ? u : s.topo[u - n].x;
  uint32_t left_edges = 0, v = u;
  uint32_t p = u < n ? gbl(s.parent_leaf)[u] : s.topo[u - n].w;
  for (uint32_t guard = 0; p != kNoParent && guard < n; ++guard) {  // depth < n
    const uint32_t j = p - n;
    if (j >= n - 1) break;
    const uint4 tj = s.topo[j];
    if (left_child(tj.z, n) == v) ++left_edges;
    v = p;
    p = tj.w;
  }
  return 2u * lo + left_edges;
}

__device__ __forceinline__ void flat_store3(float* a, uint32_t q, float x, float y, float z) {
  a[3 * (size_t)q] = x;
  a[3 * (size_t)q + 1] = y;
  a[3 * (size_t)q + 2] = z;
}

// One thread per sorted leaf k places leaf k and internal node k (k < n - 1). Geometry is
// the f32 origin-relative aggregate: centre and half extent of the AABB, com, mass and
// delta = |com - centre|; a leaf's centre is its position and its half extent zero.
__global__ void __launch_bounds__(kBlock)
k_flatten(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
          FlatOut o) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const uint32_t n = s.n;
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const uint32_t k = tl.start + e;
      {
        const uint32_t q = flat_slot(s, k);
        const float4 lf = s.leaf[k];
        o.next[q] = q + 1u;
        o.body_start[q] = k;
        o.body_count[q] = 1u;
        flat_store3(o.center, q, lf.x, lf.y, lf.z);
        flat_store3(o.half, q, 0.f, 0.f, 0.f);
        flat_store3(o.com, q, lf.x, lf.y, lf.z);
        o.mass[q] = lf.w;
        o.delta[q] = 0.f;
      }
      if (k + 1u < n) {
        const uint32_t q = flat_slot(s, n + k);
        const float4 na = s.node[k].a, nb = s.box[k].lo, nc = s.box[k].hi;
        const uint32_t lo = s.topo[k].x, hi = s.topo[k].y;
        const float cx = (nb.x + nc.x) * 0.5f, cy = (nb.y + nc.y) * 0.5f,
                    cz = (nb.z + nc.z) * 0.5f;
        const float dx = na.x - cx, dy = na.y - cy, dz = na.z - cz;
        o.next[q] = q + 2u * (hi - lo + 1u) - 1u;  // 2 leaves - 1 nodes in the subtree
        o.body_start[q] = 0u;
        o.body_count[q] = 0u;
        flat_store3(o.center, q, cx, cy, cz);
        flat_store3(o.half, q, (nc.x - nb.x) * 0.5f, (nc.y - nb.y) * 0.5f,
                    (nc.z - nb.z) * 0.5f);
        flat_store3(o.com, q, na.x, na.y, na.z);
        o.mass[q] = na.w;
        o.delta[q] = sqrtf(dx * dx + dy * dy + dz * dz);
      }
    }
  }
}

// ----------------------------------------------------------------------------------
// 7. walks. Accelerations are accumulated in origin-relative f32 coordinates with
//    optional Kahan compensation; the epilogue scatters a (x G) to the original
//    index and, in resident mode, applies the closing half-kick
//    (core/src/integrator.rs:70-73).
// ----------------------------------------------------------------------------------

struct Acc3 {
  float x, y, z, cx, cy, cz;
};

template <bool KAHAN>
__device__ __forceinline__ void acc_add(Acc3& a, float tx, float ty, float tz) {
  if (KAHAN) {
    float y, t;
    y = tx - a.cx; t = a.x + y; a.cx = (t - a.x) - y; a.x = t;
    y = ty - a.cy; t = a.y + y; a.cy = (t - a.y) - y; a.y = t;
    y = tz - a.cz; t = a.z + y; a.cz = (t - a.z) - y; a.z = t;
  } else {
    a.x += tx;
    a.y += ty;
    a.z += tz;
  }
}

template <bool KAHAN>
__device__ __forceinline__ void store_acc(const SegDesc& s, uint32_t k, const Acc3& a, int mode) {
  const uint32_t n = s.n;
  const float g = s.g;
  const float ax = (a.x - (KAHAN ? a.cx : 0.f)) * g;
  const float ay = (a.y - (KAHAN ? a.cy : 0.f)) * g;
  const float az = (a.z - (KAHAN ? a.cz : 0.f)) * g;
  const size_t i = min(s.vals[0][k], n - 1);
  if (s.acc_out) {
    s.acc_out[3 * i] = (double)ax;
    s.acc_out[3 * i + 1] = (double)ay;
    s.acc_out[3 * i + 2] = (double)az;
  } else {
    s.acc[3 * i] = ax;
    s.acc[3 * i + 1] = ay;
    s.acc[3 * i + 2] = az;
    if (mode == 1) {
      const float h = s.half_dt;
      s.v[3 * i] += ax * h;
      s.v[3 * i + 1] += ay * h;
      s.v[3 * i + 2] += az * h;
    }
  }
}

// Per-target walk (solvers/src/lbvh.rs:469-524): stackless skip-pointer traversal in
// DFS pre-order (left child first; after a subtree, its skip pointer), reference
// acceptance test, Plummer-softened monopole, self excluded by leaf index.
template <bool KAHAN, bool COUNT>
__device__ __forceinline__ void lane_walk(const SegDesc& s, uint32_t k, float4 me, Acc3& a,
                                          unsigned long long& n_int,
                                          unsigned long long& n_vis, bool quad_records,
                                          bool boxes) {
  // no contraction: the same bits whichever kernel inlines it (the per-target walk, or the
  // group walk's fallback in either record layout)
#pragma clang fp contract(off)
  const uint32_t n = s.n;
  const float eps2 = s.eps2;
  const float4* __restrict__ leaf = s.leaf;
  const Node* __restrict__ node = s.node;
  const uint32_t* __restrict__ rs = s.rs;
  uint32_t u = (n == 1) ? 0u : n;
  // each of the 2n-1 nodes is visited at most once in pre-order; the bound only
  // guards corrupted trees
  for (uint32_t guard = 0; u < 2 * n - 1 && guard < 2 * n; ++guard) {
    if (COUNT) ++n_vis;
    if (u < n) {
      const float4 p = leaf[u];
      const uint32_t nxt = (u == n - 1) ? kEnd : rs[u];
      if (u != k && p.w > 0.0f) {
        const float dx = p.x - me.x, dy = p.y - me.y, dz = p.z - me.z;
        const float r2 = dx * dx + dy * dy + dz * dz + eps2;
        const float ir = rsqrtf(r2);
        const float w = p.w * ir * ir * ir;
        acc_add<KAHAN>(a, dx * w, dy * w, dz * w);
        if (COUNT) ++n_int;
      }
      u = nxt;
    } else {
      const Node* nd = node + (u - n);
      const float4 A = nd->a, H = nd->h;
      // quadrupole builds keep classifier radii in h.zw: the skip pointer is rs[hi], with
      // hi from the topology record (loaded here only then: the group walk's rare fallback)
      uint32_t nxt = __float_as_uint(H.z);
      if (quad_records) {
        const uint32_t hi = s.topo[u - n].y;
        nxt = hi == n - 1 ? kEnd : rs[hi];
      }
      if (!(A.w > 0.0f)) {
        u = nxt;
        continue;
      }
      const float dx = A.x - me.x, dy = A.y - me.y, dz = A.z - me.z;
      const float d2 = dx * dx + dy * dy + dz * dz;
      // the box is read only for a node far enough to be accepted (the inside test). A
      // build without every box record has theta <= 1, where a target inside a node's box
      // is never accepted: it lies within sqrt(3) s / 2 + delta < s / theta + delta of the com
      bool accept = d2 >= H.x;
      if (accept && boxes) {
        const float4 B = s.box[u - n].lo, Cc = s.box[u - n].hi;
        accept = !((B.x <= me.x) & (me.x <= Cc.x) & (B.y <= me.y) & (me.y <= Cc.y) &
                   (B.z <= me.z) & (me.z <= Cc.z));
      }
      if (accept) {
        const float r2 = d2 + eps2;
        const float ir = rsqrtf(r2);
        const float w = A.w * ir * ir * ir;
        acc_add<KAHAN>(a, dx * w, dy * w, dz * w);
        if (COUNT) ++n_int;
        u = nxt;
      } else {
        const uint32_t e = left_entry(__float_as_uint(H.y));
        u = (e & kEntryNode) ? n + (e & ~kEntryNode) : e;
      }
    }
  }
}

template <bool KAHAN, bool COUNT>
__global__ void __launch_bounds__(kBlock)
k_walk_lane(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
            int mode) {
  for (uint32_t t = xcd_block<kXcdLaneRun>(); t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    if (tl.start == 0 && threadIdx.x < 6) bbox_neutral(s.bbox, threadIdx.x);
    if (threadIdx.x >= tl.count) continue;
    const uint32_t k = tl.start + threadIdx.x;
    if (s.active && !s.active[min(s.vals[0][k], s.n - 1)]) continue;
    Acc3 a = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    unsigned long long n_int = 0, n_vis = 0;
    lane_walk<KAHAN, COUNT>(s, k, s.leaf[k], a, n_int, n_vis, false, true);
    store_acc<KAHAN>(s, k, a, mode);
    if (COUNT) {
      atomicAdd(&s.counters[kCntInteractions], n_int);
      atomicAdd(&s.counters[kCntVisits], n_vis);
    }
  }
}

// Group walk: one wave walks the tree for 64 Morton-consecutive targets. Nodes to
// test are popped 64 at a time from an LDS stack (one node per lane); a node is
// accepted for the whole group when its centre of mass is at least the reference
// acceptance radius (s/theta + delta) from every point of the group's bounding box
// and the group box does not overlap the node box — so each target's own reference
// test (solvers/src/lbvh.rs:506-514) holds for every accepted node. Accepted nodes
// and reached leaves are appended to LDS interaction lists which the 64 lanes then
// evaluate against their own targets (LDS broadcast reads, no divergence). Opened
// nodes push both children. If the stack would overflow, the group falls back to
// the per-target walk.
constexpr int kStackCap = 896;  // the deepest stack measured is 520 entries
constexpr int kListCap = 64;   // lists are drained every round, which adds at most 64
constexpr int kRangeCap = 64;
constexpr int kFarCap = 128;  // far nodes are drained 64 at a time, one per lane

// Floor on the softened r^2 of a point term: a source at the target's position (dx = 0
// exactly, the target itself included) then adds 0 * finite rather than 0 * inf when the
// softening is 0. It is far below any separation f32 positions resolve at unit scale.
constexpr float kR2Floor = 1e-20f;

// Point-mass term of one source on the lane's target: a += m d / (|d|^2 + eps^2)^1.5.
// The raw v_rsq (1 ulp) serves as rsqrt: r^2 is never denormal here. FLOOR applies the
// r^2 floor; with eps^2 >= the floor, r^2 >= eps^2 already and the floor is left out.
template <bool KAHAN, bool FLOOR = true>
__device__ __forceinline__ void point_term(Acc3& a, const float4& me, const float4& p, float eps2) {
  const float dx = p.x - me.x, dy = p.y - me.y, dz = p.z - me.z;
  const float r2s = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
  const float r2 = FLOOR ? fmaxf(r2s, kR2Floor) : r2s;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float w = p.w * (ir * ir) * ir;
  if (KAHAN) {
    acc_add<true>(a, dx * w, dy * w, dz * w);
  } else {
    a.x = fmaf(dx, w, a.x);
    a.y = fmaf(dy, w, a.y);
    a.z = fmaf(dz, w, a.z);
  }
}

// Monopole + quadrupole term of one node (com/mass P, traceless Q = {xx,yy,zz,xy}, {xz,yz})
// with the softened radius r^2 = |d|^2 + eps^2, d = com - x:
//   a = m d / r^3 - Q d / r^5 + 2.5 (d.Q.d) d / r^7
//     = r^-5 [(m r^2 + 2.5 (d.Q.d) / r^2) d - Q d]
// An accepted node is never at the target (its box is disjoint from the group's), so
// r^2 > 0.
template <bool KAHAN>
__device__ __forceinline__ void quad_term(Acc3& a, const float4& me, const float4& P,
                                          const float4& B, const float2& Cq, float eps2) {
  const float dx = P.x - me.x, dy = P.y - me.y, dz = P.z - me.z;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
  const float ir = __builtin_amdgcn_rsqf(r2), ir2 = ir * ir, ir5 = (ir2 * ir2) * ir;
  const float qx = fmaf(B.x, dx, fmaf(B.w, dy, Cq.x * dz));
  const float qy = fmaf(B.w, dx, fmaf(B.y, dy, Cq.y * dz));
  const float qz = fmaf(Cq.x, dx, fmaf(Cq.y, dy, B.z * dz));
  const float dqd = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
  const float c = fmaf(P.w, r2, 2.5f * dqd * ir2);
  const float tx = fmaf(c, dx, -qx), ty = fmaf(c, dy, -qy), tz = fmaf(c, dz, -qz);
  if (KAHAN) {
    acc_add<true>(a, ir5 * tx, ir5 * ty, ir5 * tz);
  } else {
    a.x = fmaf(ir5, tx, a.x);
    a.y = fmaf(ir5, ty, a.y);
    a.z = fmaf(ir5, tz, a.z);
  }
}

// Far field of a group: a node (com at d from the group centre, mass m, traceless Q) is
// folded into a local expansion of the softened acceleration about the centre,
// a(c + y) = A0 + A1 y + A2 y y / 2 + A3 y y y / 6 (R^2 = |d|^2 + eps^2):
//   monopole   A0 = m d / R^3,  A1_ij = m (3 d_i d_j / R^5 - delta_ij / R^3),
//              A2_ijk = m (15 d_i d_j d_k / R^7
//                          - 3 (delta_ij d_k + delta_ik d_j + delta_jk d_i) / R^5),
//              A3_ijkl = m (105 d_i d_j d_k d_l / R^9 - 15 (delta_ij d_k d_l + 5 more
//                           pairings) / R^7 + 3 (delta_ij delta_kl + 2 more) / R^5)
//   quadrupole A0 = -q / R^5 + 2.5 s d / R^7 (q = Q d, s = d.q),
//              A1_ij = Q_ij / R^5 - 5 (q_i d_j + q_j d_i) / R^7 - 2.5 s delta_ij / R^7
//                      + 17.5 s d_i d_j / R^9
// (the successive derivatives of the field in y: each order adds one factor |y| / R to
// the truncation error; the quadrupole's second-order term is dropped: it is smaller than
// the neglected fourth-order monopole term for well-separated nodes). Coefficients:
// L[0..2] = A0, L[3..8] = A1 {xx,yy,zz,xy,xz,yz},
// L[9..18] = A2 {xxx,xxy,xxz,xyy,xyz,xzz,yyy,yyz,yzz,zzz},
// L[19..33] = A3 {xxxx,xxxy,xxxz,xxyy,xxyz,xxzz,xyyy,xyyz,xyzz,xzzz,yyyy,yyyz,yyzz,yzzz,zzzz}.
constexpr int kLocal = 34;

__device__ __forceinline__ void far_m2l(float* L, float x, float y, float z, float m,
                                        const float4& Q0, const float4& Q1, float eps2) {
  const float R2 = fmaf(x, x, fmaf(y, y, fmaf(z, z, eps2)));
  const float ir = __builtin_amdgcn_rsqf(R2), ir2 = ir * ir;
  const float mr3 = m * ir * ir2, ir5 = ir * ir2 * ir2, ir7 = ir5 * ir2;
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  // monopole
  L[0] = fmaf(mr3, x, L[0]);
  L[1] = fmaf(mr3, y, L[1]);
  L[2] = fmaf(mr3, z, L[2]);
  const float t = 3.0f * m * ir5;
  L[3] += fmaf(t, xx, -mr3);
  L[4] += fmaf(t, yy, -mr3);
  L[5] += fmaf(t, zz, -mr3);
  L[6] = fmaf(t, xy, L[6]);
  L[7] = fmaf(t, xz, L[7]);
  L[8] = fmaf(t, yz, L[8]);
  const float w = 15.0f * m * ir7;
  L[9] += x * fmaf(w, xx, -3.0f * t);
  L[10] += y * fmaf(w, xx, -t);
  L[11] += z * fmaf(w, xx, -t);
  L[12] += x * fmaf(w, yy, -t);
  L[13] = fmaf(w * xy, z, L[13]);
  L[14] += x * fmaf(w, zz, -t);
  L[15] += y * fmaf(w, yy, -3.0f * t);
  L[16] += z * fmaf(w, yy, -t);
  L[17] += y * fmaf(w, zz, -t);
  L[18] += z * fmaf(w, zz, -3.0f * t);
  // third order: 105 m X / R^9 - w (sum over equal-index pairs of the other two d's)
  // + t (number of pairings into equal pairs), X the component's product of four d's
  const float a9 = 7.0f * w * ir2;  // 105 m / R^9
  L[19] += fmaf(a9 * xx, xx, fmaf(-6.0f * w, xx, 3.0f * t));
  L[20] += xy * fmaf(a9, xx, -3.0f * w);
  L[21] += xz * fmaf(a9, xx, -3.0f * w);
  L[22] += fmaf(a9 * xx, yy, fmaf(-w, xx + yy, t));
  L[23] += yz * fmaf(a9, xx, -w);
  L[24] += fmaf(a9 * xx, zz, fmaf(-w, xx + zz, t));
  L[25] += xy * fmaf(a9, yy, -3.0f * w);
  L[26] += xz * fmaf(a9, yy, -w);
  L[27] += xy * fmaf(a9, zz, -w);
  L[28] += xz * fmaf(a9, zz, -3.0f * w);
  L[29] += fmaf(a9 * yy, yy, fmaf(-6.0f * w, yy, 3.0f * t));
  L[30] += yz * fmaf(a9, yy, -3.0f * w);
  L[31] += fmaf(a9 * yy, zz, fmaf(-w, yy + zz, t));
  L[32] += yz * fmaf(a9, zz, -3.0f * w);
  L[33] += fmaf(a9 * zz, zz, fmaf(-6.0f * w, zz, 3.0f * t));
  // quadrupole {xx,yy,zz,xy} = Q0, {xz,yz} = Q1.xy
  const float qx = fmaf(Q0.x, x, fmaf(Q0.w, y, Q1.x * z));
  const float qy = fmaf(Q0.w, x, fmaf(Q0.y, y, Q1.y * z));
  const float qz = fmaf(Q1.x, x, fmaf(Q1.y, y, Q0.z * z));
  const float s = fmaf(x, qx, fmaf(y, qy, z * qz));
  const float f = 2.5f * s * ir7, e = 7.0f * f * ir2, g5 = 5.0f * ir7;
  L[0] += fmaf(f, x, -qx * ir5);
  L[1] += fmaf(f, y, -qy * ir5);
  L[2] += fmaf(f, z, -qz * ir5);
  L[3] += fmaf(Q0.x, ir5, fmaf(e, xx, -f)) - 2.0f * g5 * qx * x;
  L[4] += fmaf(Q0.y, ir5, fmaf(e, yy, -f)) - 2.0f * g5 * qy * y;
  L[5] += fmaf(Q0.z, ir5, fmaf(e, zz, -f)) - 2.0f * g5 * qz * z;
  L[6] += fmaf(Q0.w, ir5, e * xy) - g5 * fmaf(qx, y, qy * x);
  L[7] += fmaf(Q1.x, ir5, e * xz) - g5 * fmaf(qx, z, qz * x);
  L[8] += fmaf(Q1.y, ir5, e * yz) - g5 * fmaf(qy, z, qz * y);
}

// a(c + y) from the local coefficients: a_i = A0_i + A1_ij y_j + A2_ijk y_j y_k / 2
// + A3_ijkl y_j y_k y_l / 6, the symmetric sums written over distinct monomials.
__device__ __forceinline__ void far_l2p(const float* L, float x, float y, float z, float& ax,
                                        float& ay, float& az) {
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  ax = L[0] + fmaf(L[3], x, fmaf(L[6], y, L[7] * z)) +
       0.5f * fmaf(L[9], xx, fmaf(L[12], yy, fmaf(L[14], zz,
              2.0f * fmaf(L[10], xy, fmaf(L[11], xz, L[13] * yz)))));
  ay = L[1] + fmaf(L[6], x, fmaf(L[4], y, L[8] * z)) +
       0.5f * fmaf(L[10], xx, fmaf(L[15], yy, fmaf(L[17], zz,
              2.0f * fmaf(L[12], xy, fmaf(L[13], xz, L[16] * yz)))));
  az = L[2] + fmaf(L[7], x, fmaf(L[8], y, L[5] * z)) +
       0.5f * fmaf(L[11], xx, fmaf(L[16], yy, fmaf(L[18], zz,
              2.0f * fmaf(L[13], xy, fmaf(L[14], xz, L[17] * yz)))));
  // third order over the cubic monomials scaled by 1 / (p! q! r!)
  const float c0 = xx * x * (1.0f / 6.0f), c1 = 0.5f * xx * y, c2 = 0.5f * xx * z,
              c3 = 0.5f * x * yy, c4 = xy * z, c5 = 0.5f * x * zz, c6 = yy * y * (1.0f / 6.0f),
              c7 = 0.5f * yy * z, c8 = 0.5f * y * zz, c9 = zz * z * (1.0f / 6.0f);
  ax += fmaf(L[19], c0, fmaf(L[20], c1, fmaf(L[21], c2, fmaf(L[22], c3, fmaf(L[23], c4,
        fmaf(L[24], c5, fmaf(L[25], c6, fmaf(L[26], c7, fmaf(L[27], c8, L[28] * c9)))))))));
  ay += fmaf(L[20], c0, fmaf(L[22], c1, fmaf(L[23], c2, fmaf(L[25], c3, fmaf(L[26], c4,
        fmaf(L[27], c5, fmaf(L[29], c6, fmaf(L[30], c7, fmaf(L[31], c8, L[32] * c9)))))))));
  az += fmaf(L[21], c0, fmaf(L[23], c1, fmaf(L[24], c2, fmaf(L[26], c3, fmaf(L[27], c4,
        fmaf(L[28], c5, fmaf(L[30], c6, fmaf(L[31], c7, fmaf(L[32], c8, L[33] * c9)))))))));
}

// The queued far nodes (node indices) into this lane's share of the local expansion:
// lane e takes entries e, e + 64, ...
__device__ __forceinline__ void far_drain(float* L, const SegDesc& s, const uint32_t* fl,
                                          uint32_t nf, uint32_t lane, float cx, float cy,
                                          float cz, float eps2) {
  for (uint32_t e = lane; e < nf; e += kGroup) {
    const uint32_t j = fl[e];
    const float4 A = gload4(reinterpret_cast<const float4*>(s.node), 2 * (size_t)j);
    far_m2l(L, A.x - cx, A.y - cy, A.z - cz, A.w, gload4(s.quad, 2 * (size_t)j),
            gload4(s.quad, 2 * (size_t)j + 1), eps2);
  }
}

// Lanes below this one with their bit set in the wave mask m (v_mbcnt_lo/hi: two VALU
// instructions, against an AND and two bit counts on a per-lane mask).
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ float wave_sumf(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Bodies of the buckets in `rl[0, nr)` (first leaf, count) on the lane's target. The
// range is wave-uniform, so the body records come through scalar loads (the leaf array
// is read-only during the walk: a constant-address-space pointer gives s_load).
template <bool KAHAN>
__device__ __forceinline__ uint32_t bucket_terms(Acc3& a, const float4& me, const SegDesc& s,
                                                 const uint2* rl, uint32_t nr, float eps2) {
  const auto* leaf = (const __attribute__((address_space(4))) float*)s.leaf;
  uint32_t total = 0;
  for (uint32_t r = 0; r < nr; ++r) {
    const uint2 rg = rl[r];
    const uint32_t lo = __builtin_amdgcn_readfirstlane(rg.x);
    const uint32_t c = __builtin_amdgcn_readfirstlane(rg.y);
    for (uint32_t e = 0; e < c; ++e) {
      const auto* q = leaf + 4 * (size_t)(lo + e);
      point_term<KAHAN>(a, me, make_float4(q[0], q[1], q[2], q[3]), eps2);
    }
    total += c;
  }
  return total;
}

// The far-field build holds a 34-coefficient local expansion per lane: without a bound
// the compiler takes about 147 VGPRs (3 waves per SIMD: walk +21%); at 4 waves per SIMD it
// fits 128 VGPRs without spilling. The counting build is left unbounded (it only counts).
template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

template <bool KAHAN, bool COUNT, bool QUAD, bool FAR, bool WIDE>
__global__ void __launch_bounds__(kGroup)
__attribute__((amdgpu_waves_per_eu(COUNT ? 1 : 4)))
k_walk_group(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
             int mode) {
  constexpr int QN = QUAD ? kListCap : 1;
  __shared__ uint32_t stk[kStackCap];
  __shared__ float4 pl[kListCap];
  __shared__ float4 qa[QN], qb[QN];
  __shared__ float2 qc[QN];
  __shared__ uint2 rl[kRangeCap];
  __shared__ uint32_t fl[FAR ? kFarCap : 1];
  const uint32_t lane = threadIdx.x;
  for (uint32_t t = xcd_block(); t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    if (tl.start == 0 && lane < 6) bbox_neutral(s.bbox, lane);
    const uint32_t n = s.n;
    const float eps2 = s.eps2;
    const bool soft = eps2 >= kR2Floor;  // r^2 >= eps^2 needs no floor
    const uint32_t bucket = s.leaf_size;
    // the disjoint-box test is needed only above theta = 1 (theta >= 0: compared as bits,
    // a scalar test)
    const bool loose_theta = __float_as_uint(s.theta) > __float_as_uint(1.0f);
    const bool need_box = loose_theta || bucket > 1u;  // box records (bounds, leaf range)
    const uint64_t leaf_base = (uint64_t)s.leaf, node_base = (uint64_t)s.node;
    // The compact form addresses leaves, classifier records and quadrupoles by 32-bit
    // byte offsets from the leaf array (the launch takes the wide form when a segment's
    // arrays do not fit; a segment that does not fit here is walked per target)
    const bool fits = WIDE || walk_off32(leaf_base, node_base, (uint64_t)s.quad, n);
    const char* const lbase = reinterpret_cast<const char*>(s.leaf);
    const char* const lbase16 = lbase + 16;  // second float4 of a record
    // stack entries before a group falls back to the per-target walk
    const uint32_t scap = s.stack_cap && s.stack_cap < (uint32_t)kStackCap ? s.stack_cap
                                                                           : (uint32_t)kStackCap;
    const uint32_t dn32 = (uint32_t)(node_base - leaf_base),
                   dq32 = (uint32_t)((uint64_t)s.quad - leaf_base);
    // the walk groups (subtrees of at most kGroup bodies, marked by the tree build)
    // that start in this chunk of sorted leaves, one after the other
    const uint32_t ge = lane < tl.count ? gbl(s.gend)[tl.start + lane] : kEnd;
    unsigned long long starts = __ballot(ge != kEnd);
    unsigned long long t_int = 0, t_vis = 0, t_quad = 0, t_far = 0, t_rounds = 0;
    uint32_t t_peak = 0;
    while (starts != 0ull) {
      const uint32_t l0 = (uint32_t)__builtin_ctzll(starts);
      starts &= starts - 1ull;
      const uint32_t g0 = tl.start + l0;
      const uint32_t gl = (uint32_t)__shfl((int)ge, (int)l0);
      const uint32_t gcount = min(min(gl - g0 + 1u, (uint32_t)kGroup), n - g0);
      // A group of at most 32 bodies gives each body rep lanes (2, 3 or 4): the lists are
      // drained with stride rep, and the partial sums are folded into the body's first
      // lane (its owner) before the store.
      // (at most group_lanes; rep >= 1 also lets the compiler count the drains' trips)
      const uint32_t rep = max(min(gcount > 32u ? 1u : (gcount > 21u ? 2u : (gcount > 16u ? 3u : 4u)),
                                   s.group_lanes), 1u);
      const uint32_t sub = (uint32_t)(lane >= gcount) + (uint32_t)(lane >= 2u * gcount) +
                           (uint32_t)(lane >= 3u * gcount);
      const bool owner = lane < gcount;
      const bool valid = lane < rep * gcount;
      const uint32_t k = g0 + (valid ? lane - sub * gcount : 0u);
      const uint32_t e0 = valid ? sub : 0xFFFFu;  // this lane's first list entry
      if (s.active) {  // masked walk: a group without an active body is skipped
        const bool on = owner && s.active[min(s.vals[0][k], n - 1)] != 0;
        if (__ballot(on) == 0ull) continue;
      }
      const float4 me = gload4(s.leaf, k);
      const float gx0 = wave_minf(me.x), gy0 = wave_minf(me.y), gz0 = wave_minf(me.z);
      const float gx1 = wave_maxf(me.x), gy1 = wave_maxf(me.y), gz1 = wave_maxf(me.z);
      // far field: group centre and radius; this lane's share of the local expansion
      const float cx = 0.5f * (gx0 + gx1), cy = 0.5f * (gy0 + gy1), cz = 0.5f * (gz0 + gz1);
      const float hx = 0.5f * (gx1 - gx0), hy = 0.5f * (gy1 - gy0), hz = 0.5f * (gz1 - gz0);
      const float rg = FAR ? sqrtf(fmaf(hx, hx, fmaf(hy, hy, hz * hz))) : 0.0f;
      float loc[FAR ? kLocal : 1];
#pragma unroll
      for (int c = 0; c < (FAR ? kLocal : 1); ++c) loc[c] = 0.0f;
      Acc3 a = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      unsigned long long n_int = 0, n_vis = 0, n_quad = 0, n_far = 0, n_rounds = 0;
      uint32_t sp = fits ? 1u : 0u, np = 0, nq = 0, nr = 0, nf = 0, peak = 1;
      bool overflow = !fits;
      // stack entries: leaf g -> g, internal node g -> g | 2^31 (the root is node 0)
      if (lane == 0) stk[0] = (n == 1) ? 0u : kEntryNode;
      __syncthreads();
      // The rounds, specialised on whether the node tests need the box record (wave-
      // uniform per segment), so that the common case carries none of its instructions
      auto rounds = [&](auto box_tag) {
        constexpr bool BOX = decltype(box_tag)::value;
        // every node is pushed at most once, so a round count above 2n only happens on a
        // corrupted tree; the bound keeps such input from spinning forever
        for (uint32_t round = 0; sp > 0; ++round) {
          if (round > 2 * n + 2) {
            overflow = true;
            break;
          }
          const uint32_t m = sp < 64u ? sp : 64u;
          const uint32_t base = sp - m;
          const bool live = lane < m;
          const uint32_t u = live ? stk[base + lane] : 0u;  // beyond the round: leaf 0, not live
          sp = base;
          if (COUNT && lane == 0) {
            n_vis += m;
            ++n_rounds;
          }
          // Classify without divergent branches: every lane loads one record and the node
          // fields, and the predicates below are balloted, so the ballots that build the
          // lists always see the full wave.
          // Lists: point/monopole, quadrupole node, open (children pushed), bucket (a node of
          // at most leaf_size bodies that is not accepted: its bodies are summed directly),
          // far (local expansion).
          const bool is_node = live & ((int32_t)u < 0);
          const bool is_leaf = live & ((int32_t)u >= 0);
          // node index (0 on a leaf lane, which reads node 0's record, a line every such
          // lane shares, and ignores it) and leaf index
          const uint32_t ni = is_node ? u & ~kEntryNode : 0u, li = u;
          const size_t nr2 = 2 * (size_t)ni;  // record as two float4
          // a leaf's record (position, mass) has the layout of a node's first float4 (com,
          // mass), so one load serves both kinds
          float4 A, H;
          if (!WIDE) {
            const uint32_t noff = dn32 + (ni << 5);
            A = gload4_at(lbase, is_leaf ? li << 4 : noff);
            H = gload4_at(lbase16, noff);
          } else {
            const uint64_t abase = is_leaf ? leaf_base : node_base;
            A = gload4(reinterpret_cast<const float4*>(abase), is_leaf ? li : 2u * ni);
            H = gload4(reinterpret_cast<const float4*>(s.node), nr2 + 1);
          }
          // The box record only where it is needed (wave-uniform): the disjoint-box test
          // (apart) and the bucket test's leaf range
          bool apart = true;
          uint32_t nlo = 0u, ncount = 2u;  // an internal node holds at least two bodies
          if (BOX) {
            const float4* brec = reinterpret_cast<const float4*>(s.box);
            const float4 B = gload4(brec, nr2), Cc = gload4(brec, nr2 + 1);
            // Every point of the node's box lies within sqrt(3) s / 2 + delta of its com
            // (half its diagonal from the box centre, delta from there to the com), and an
            // accepted com is at least s / theta + delta from the group box, so for theta <= 1
            // (1 / theta > sqrt(3) / 2 with a margin) acceptance implies disjoint boxes. (A
            // node of size 0 is one point, whose monopole is exact for every body.)
            if (loose_theta)
              apart = (Cc.x < gx0) | (B.x > gx1) | (Cc.y < gy0) | (B.y > gy1) | (Cc.z < gz0) |
                      (B.z > gz1);
            nlo = __float_as_uint(B.w);
            ncount = __float_as_uint(Cc.w) - nlo + 1u;
          }
          // com relative to the group centre; distance from the group box to the com
          const float dx = A.x - cx, dy = A.y - cy, dz = A.z - cz;
          const float ex = fmaxf(fabsf(dx) - hx, 0.f);
          const float ey = fmaxf(fabsf(dy) - hy, 0.f);
          const float ez = fmaxf(fabsf(dz) - hz, 0.f);
          const float dmin2 = ex * ex + ey * ey + ez * ez;
          const bool massive = A.w > 0.0f;
          const bool node_live = is_node & massive;
          const bool accept = node_live & apart & (dmin2 >= H.x);  // boxes disjoint too
          const bool near = node_live & !accept;
          const bool bkt = BOX & near & (ncount <= bucket);  // (leaf_size > 1 needs boxes)
          const bool open = near & !bkt;
          // accepted and well separated from the whole group -> local expansion (node radius
          // about its com from the record)
          bool far = false;
          if (FAR) {
            const float rn = H.w + rg;
            far = accept & (rn * rn <= s.far2 * fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
          }
          // an accepted node that also passes the opening test at theta * mono_ratio is a
          // monopole: the squared band radius is in the record
          const bool mono = QUAD && accept && dmin2 >= H.z;
          const bool point = (is_leaf & massive) | (!QUAD & accept) | (mono & !far);
          const bool quadn = QUAD & accept & !point & !far;
          const uint32_t cl = left_entry(__float_as_uint(H.y)),
                         cr = right_entry(__float_as_uint(H.y));
          const unsigned long long m1 = __ballot(point), m2 = __ballot(quadn),
                                   m3 = __ballot(open);
          unsigned long long m4 = 0ull;
          if (BOX) m4 = __ballot(bkt);
          const uint32_t c3 = (uint32_t)__popcll(m3);
          if (sp + 2u * c3 > scap) {
            overflow = true;
            break;
          }
          if (point) pl[np + lanes_below(m1)] = A;
          if (QUAD && quadn) {  // only accepted nodes fetch their moments
            const uint32_t q = nq + lanes_below(m2);
            float4 q0, q1;
            if (!WIDE) {
              const uint32_t qoff = dq32 + (ni << 5);
              q0 = gload4_at(lbase, qoff);
              q1 = gload4_at(lbase16, qoff);
            } else {
              q0 = gload4(s.quad, 2 * (size_t)ni);
              q1 = gload4(s.quad, 2 * (size_t)ni + 1);
            }
            qa[q] = A;
            qb[q] = q0;
            qc[q] = make_float2(q1.x, q1.y);
          }
          if (open) {
            const uint32_t q = sp + 2u * lanes_below(m3);
            stk[q] = cr;
            stk[q + 1] = cl;  // left child popped first
          }
          if (bkt) rl[nr + lanes_below(m4)] = make_uint2(nlo, ncount);
          if (FAR) {
            const unsigned long long m5 = __ballot(far);
            if (far) fl[nf + lanes_below(m5)] = ni;
            nf += (uint32_t)__popcll(m5);
          }
          np += (uint32_t)__popcll(m1);
          nq += (uint32_t)__popcll(m2);
          nr += (uint32_t)__popcll(m4);
          sp += 2u * c3;
          peak = sp > peak ? sp : peak;
          __syncthreads();
          // The drains: with one lane per body (groups of more than 32) every lane takes
          // every entry, so the loop runs on a wave-uniform index (scalar loop control,
          // no per-lane bounds); the lanes beyond the group are masked off (their terms
          // would never be stored; masked, the walk measured 1.4% faster at the same
          // instruction count).
          if (np != 0u) {
            if (rep == 1u && soft) {
              if (valid)
  #pragma unroll 2
              for (uint32_t e = 0; e < np; ++e) point_term<KAHAN, false>(a, me, pl[e], eps2);
            } else if (rep == 1u) {
              if (valid)
  #pragma unroll 2
              for (uint32_t e = 0; e < np; ++e) point_term<KAHAN>(a, me, pl[e], eps2);
            } else {
  #pragma unroll 2
              for (uint32_t e = e0; e < np; e += rep) point_term<KAHAN>(a, me, pl[e], eps2);
            }
            if (COUNT && e0 < np) n_int += (np - e0 + rep - 1u) / rep;
            np = 0;
          }
          if (nr != 0u) {  // bucket bodies come through scalar loads: owners only
            if (owner) {
              const uint32_t c = bucket_terms<KAHAN>(a, me, s, rl, nr, eps2);
              if (COUNT) n_int += c;
            }
            nr = 0;
          }
          if (QUAD && nq != 0u) {
            if (rep == 1u) {
              if (valid)
  #pragma unroll 2
              for (uint32_t e = 0; e < nq; ++e) quad_term<KAHAN>(a, me, qa[e], qb[e], qc[e], eps2);
            } else {
  #pragma unroll 2
              for (uint32_t e = e0; e < nq; e += rep)
                quad_term<KAHAN>(a, me, qa[e], qb[e], qc[e], eps2);
            }
            if (COUNT && e0 < nq) n_quad += (nq - e0 + rep - 1u) / rep;
            nq = 0;
          }
          if (FAR && nf >= 64u) {  // one far node per lane
            far_drain(loc, s, fl, nf, lane, cx, cy, cz, eps2);
            if (COUNT && lane == 0) n_far += nf;
            nf = 0;
          }
          __syncthreads();
        }
      };
      if (need_box) rounds(BoolTag<true>{});
      else rounds(BoolTag<false>{});
      if (overflow) {
        a = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        n_int = 0;
        n_quad = 0;
        if (owner) lane_walk<KAHAN, COUNT>(s, k, me, a, n_int, n_vis, QUAD, need_box);
        if (lane == 0) atomicAdd(&s.counters[kCntOverflow], 1ull);
      }  // otherwise the lists were drained in the last round
      if (rep > 1u) {  // fold each body's lanes into its owner (lanes lane + j gcount < 64)
        const float px = a.x - (KAHAN ? a.cx : 0.f), py = a.y - (KAHAN ? a.cy : 0.f),
                    pz = a.z - (KAHAN ? a.cz : 0.f);
        for (uint32_t j = 1; j < rep; ++j) {
          const int src = (int)((lane + j * gcount) & 63u);
          const float qx = __shfl(px, src), qy = __shfl(py, src), qz = __shfl(pz, src);
          if (owner) acc_add<KAHAN>(a, qx, qy, qz);
        }
      }
      if (FAR && !overflow) {
        far_drain(loc, s, fl, nf, lane, cx, cy, cz, eps2);
        if (COUNT && lane == 0) n_far += nf;
#pragma unroll
        for (int c = 0; c < (FAR ? kLocal : 1); ++c) loc[c] = wave_sumf(loc[c]);
        if (owner) {  // the other lanes' partial sums are folded already
          float fx, fy, fz;
          far_l2p(loc, me.x - cx, me.y - cy, me.z - cz, fx, fy, fz);
          acc_add<KAHAN>(a, fx, fy, fz);
        }
      }
      if (owner) store_acc<KAHAN>(s, k, a, mode);
      if (COUNT) {
        // n_int is per target (uniform over valid lanes unless the group fell back)
        t_int += wave_sum_u64(valid ? n_int : 0ull);
        t_quad += wave_sum_u64(valid ? n_quad : 0ull);
        t_vis += wave_sum_u64(n_vis);
        t_far += wave_sum_u64(n_far);
        t_rounds += wave_sum_u64(n_rounds);
        t_peak = peak > t_peak ? peak : t_peak;
      }
      __syncthreads();
    }
    if (COUNT && lane == 0 && (t_int | t_quad | t_vis)) {
      atomicAdd(&s.counters[kCntInteractions], t_int);
      atomicAdd(&s.counters[kCntQuad], t_quad);
      atomicAdd(&s.counters[kCntVisits], t_vis);
      atomicMax(&s.counters[kCntMaxStack], (unsigned long long)t_peak);
      atomicAdd(&s.counters[kCntFar], t_far);
      atomicAdd(&s.counters[kCntRounds], t_rounds);
    }
  }
}

// ----------------------------------------------------------------------------------
// 8. FP64 diagnostics (core/src/diagnostics.rs:8-68): per-tile partial sums in a
//    fixed reduction order, then a fixed-order sum over each segment's tiles.
//    Partial record: KE, P xyz, L xyz, sum m*x xyz, sum m  (11 doubles).
// ----------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_diag_partial(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
               double* __restrict__ partial) {
  __shared__ double sh[kDiagRecord][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    double acc[kDiagRecord];
    for (int q = 0; q < kDiagRecord; ++q) acc[q] = 0.0;
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const size_t i = tl.start + e;
      const double m = s.m[i];
      const double x = s.x[3 * i], y = s.x[3 * i + 1], z = s.x[3 * i + 2];
      const double vx = s.v[3 * i], vy = s.v[3 * i + 1], vz = s.v[3 * i + 2];
      acc[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
      acc[1] += vx * m;
      acc[2] += vy * m;
      acc[3] += vz * m;
      acc[4] += (y * vz - z * vy) * m;
      acc[5] += (z * vx - x * vz) * m;
      acc[6] += (x * vy - y * vx) * m;
      acc[7] += x * m;
      acc[8] += y * m;
      acc[9] += z * m;
      acc[10] += m;
    }
    for (int q = 0; q < kDiagRecord; ++q) {
      const double r = wave_sum(acc[q]);
      if (lane == 0) sh[q][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < kDiagRecord) {
      double r = 0.0;
      for (int k = 0; k < kBlock / 64; ++k) r += sh[threadIdx.x][k];
      partial[(size_t)t * kDiagRecord + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

__global__ void k_diag_final(const SegDesc* __restrict__ segs, uint32_t nseg,
                             const double* __restrict__ partial, double* __restrict__ out) {
  const uint32_t sidx = blockIdx.x;
  if (sidx >= nseg || threadIdx.x >= kDiagRecord) return;
  const SegDesc& s = segs[sidx];
  double r = 0.0;
  for (uint32_t t = s.sort_tile0; t < s.sort_tile0 + s.sort_tiles; ++t)
    r += partial[(size_t)t * kDiagRecord + threadIdx.x];
  out[(size_t)sidx * kDiagRecord + threadIdx.x] = r;
}

// ----------------------------------------------------------------------------------
// 9. FP64 direct summation (solvers/src/direct_sum.rs:25-45 pair kernel) onto a
//    target subset, and the FP64 softened potential (solvers/src/potential.rs:17-27)
//    as per-block partials. Sources are staged through LDS.
// ----------------------------------------------------------------------------------

constexpr int kDsTile = 256;

// 1/sqrt(x) in f64: the f32 hardware estimate (relative error < 2^-22) refined by two
// Newton steps y <- y (3/2 - x y^2 / 2), each of which squares the relative error, so the
// result is within a few f64 ulps. Outside the range where x is a normal f32 it falls back
// to the IEEE square root and division.
__device__ __forceinline__ double rsqrt_f64(double x) {
  if (__builtin_expect(!(x > 1e-36 && x < 1e36), 0)) return 1.0 / sqrt(x);
  double y = (double)__builtin_amdgcn_rsqf((float)x);
  const double hx = 0.5 * x;
  y = y * fma(-hx * y, y, 1.5);
  y = y * fma(-hx * y, y, 1.5);
  return y;
}

// Targets [blockIdx.x * 256, +256) against the sources of chunk blockIdx.y, [c * len,
// min(n, (c + 1) * len)), self excluded. With one chunk the sums go straight to out
// (3 per target); with several, chunk c writes partial[(c * nt + t) * 3 + axis] and
// k_direct_reduce adds the chunks in index order, so results do not depend on timing.
__global__ void __launch_bounds__(kDsTile)
k_direct_f64(const double* __restrict__ x, const double* __restrict__ m, uint32_t n,
             const uint32_t* __restrict__ targets, uint32_t nt, double g, double eps2,
             uint32_t chunk_len, double* __restrict__ out) {
  __shared__ double sx[kDsTile], sy[kDsTile], sz[kDsTile], sm[kDsTile];
  const uint32_t t = blockIdx.x * kDsTile + threadIdx.x;
  const bool act = t < nt;
  const uint32_t i = act ? targets[t] : 0u;
  const double xi = act ? x[3 * (size_t)i] : 0.0, yi = act ? x[3 * (size_t)i + 1] : 0.0,
               zi = act ? x[3 * (size_t)i + 2] : 0.0;
  const uint32_t j0 = blockIdx.y * chunk_len;
  const uint32_t j1 = (n - j0 < chunk_len) ? n : j0 + chunk_len;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (uint32_t base = j0; base < j1; base += kDsTile) {
    const uint32_t j = base + threadIdx.x;
    if (j < j1) {
      sx[threadIdx.x] = x[3 * (size_t)j];
      sy[threadIdx.x] = x[3 * (size_t)j + 1];
      sz[threadIdx.x] = x[3 * (size_t)j + 2];
      sm[threadIdx.x] = m[j];
    }
    __syncthreads();
    const uint32_t cnt = (j1 - base < (uint32_t)kDsTile) ? j1 - base : (uint32_t)kDsTile;
    for (uint32_t q = 0; q < cnt; ++q) {
      const double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double ir = rsqrt_f64(r2);
      const double w = (base + q == i) ? 0.0 : sm[q] * (ir * ir * ir);
      ax += dx * w;
      ay += dy * w;
      az += dz * w;
    }
    __syncthreads();
  }
  if (act) {
    double* o = out + ((size_t)blockIdx.y * nt + t) * 3;
    o[0] = g * ax;
    o[1] = g * ay;
    o[2] = g * az;
  }
}

// out[3 t + axis] = sum over chunks c = 0, 1, ... of partial[(c * nt + t) * 3 + axis].
__global__ void __launch_bounds__(kDsTile)
k_direct_reduce(const double* __restrict__ partial, uint32_t nt, uint32_t chunks,
                double* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * kDsTile + threadIdx.x;
  if (e >= 3 * (size_t)nt) return;
  double r = 0.0;
  for (uint32_t c = 0; c < chunks; ++c) r += partial[(size_t)c * 3 * nt + e];
  out[e] = r;
}

// Softened potential: rows [blockIdx*256, ...) against all j > i; one partial per block.
__global__ void __launch_bounds__(kDsTile)
k_potential_f64(const double* __restrict__ x, const double* __restrict__ m, uint32_t n, double g,
                double eps2, double* __restrict__ partial) {
  __shared__ double sx[kDsTile], sy[kDsTile], sz[kDsTile], sm[kDsTile];
  __shared__ double red[kDsTile / 64];
  const uint32_t i = blockIdx.x * kDsTile + threadIdx.x;
  const bool act = i < n;
  const double xi = act ? x[3 * (size_t)i] : 0.0, yi = act ? x[3 * (size_t)i + 1] : 0.0,
               zi = act ? x[3 * (size_t)i + 2] : 0.0, mi = act ? m[i] : 0.0;
  double u = 0.0;
  for (uint32_t base = blockIdx.x * kDsTile; base < n; base += kDsTile) {
    const uint32_t j = base + threadIdx.x;
    if (j < n) {
      sx[threadIdx.x] = x[3 * (size_t)j];
      sy[threadIdx.x] = x[3 * (size_t)j + 1];
      sz[threadIdx.x] = x[3 * (size_t)j + 2];
      sm[threadIdx.x] = m[j];
    }
    __syncthreads();
    const uint32_t cnt = (n - base < (uint32_t)kDsTile) ? n - base : (uint32_t)kDsTile;
    for (uint32_t q = 0; q < cnt; ++q) {
      const double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
      const double ir = rsqrt_f64(dx * dx + dy * dy + dz * dz + eps2);
      u += (base + q > i) ? sm[q] * ir : 0.0;  // pairs j > i only (selected, not multiplied)
    }
    __syncthreads();
  }
  u = act ? -g * mi * u : 0.0;
  u = wave_sum(u);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = u;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < kDsTile / 64; ++k) r += red[k];
    partial[blockIdx.x] = r;
  }
}

__global__ void k_sum_f64(const double* __restrict__ in, uint32_t n, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double r = 0.0;
  for (uint32_t i = 0; i < n; ++i) r += in[i];
  *out = r;
}

__global__ void k_seg_reset(const SegDesc* __restrict__ segs, uint32_t nseg) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nseg * 6u) return;
  bbox_neutral(segs[q / 6u].bbox, (int)(q % 6u));
}

// ----------------------------------------------------------------------------------
// host-side launch wrappers
// ----------------------------------------------------------------------------------

static inline uint32_t grid_for(uint32_t ntiles, uint32_t cap = 65535u) {
  return ntiles == 0 ? 1u : (ntiles < cap ? ntiles : cap);
}

// Grid for the XCD-ordered kernels: a multiple of the span (idle blocks skip the tile loop).
template <uint32_t RUN = kXcdRun>
static inline uint32_t xcd_grid(uint32_t ntiles, uint32_t max_blocks = 65535u) {
  const uint32_t span = kXcds * RUN, cap = max_blocks / span * span;
  if (ntiles < span) return ntiles == 0 ? 1u : ntiles;
  const uint32_t g = (ntiles + span - 1) / span * span;
  return g < cap ? g : cap;
}

hipError_t launch_seg_reset(hipStream_t st, const SegDesc* segs, uint32_t nseg) {
  if (nseg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_reset, dim3((nseg * 6 + 255) / 256), dim3(256), 0, st, segs, nseg);
  return hipGetLastError();
}

hipError_t launch_bbox(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net) {
  hipLaunchKernelGGL(k_bbox, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_kick_drift_bbox(hipStream_t st, const SegDesc* segs, const Tile* et,
                                  uint32_t net) {
  hipLaunchKernelGGL(k_kick_drift_bbox, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_encode(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* et,
                         uint32_t net, uint32_t* lanes_out) {
  const int dst = sort_start(md.key63, md.radix_bits);
  if (md.key63)
    hipLaunchKernelGGL(k_encode<uint64_t>, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net,
                       lanes_out, dst);
  else
    hipLaunchKernelGGL(k_encode<uint32_t>, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net,
                       lanes_out, dst);
  return hipGetLastError();
}

hipError_t launch_bounds_only(hipStream_t st, const Mode& md, const SegDesc* segs, uint32_t nseg) {
  hipLaunchKernelGGL(k_bounds, dim3((nseg + 63) / 64), dim3(64), 0, st, segs, nseg,
                     md.key63 ? 21 : 10);
  return hipGetLastError();
}

hipError_t launch_iota(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* et,
                       uint32_t net) {
  hipLaunchKernelGGL(k_iota, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net,
                     sort_start(md.key63, md.radix_bits));
  return hipGetLastError();
}

// The keys start in buffer sort_start() and end in buffer 0.
template <class K, int B>
static void sort_passes(hipStream_t st, const SegDesc* segs, uint32_t nseg, const Tile* et,
                        uint32_t net, const SortScratch& sc) {
  constexpr bool key63 = sizeof(K) == 8;
  uint32_t* tile_hist = sc.tile_hist;
  uint32_t* tile_off = sc.tile_off;
  int src = sort_start(key63, B);
  for (int pass = 0; pass < sort_passes_for(key63, B); ++pass) {
    const int shift = pass * B;
    hipLaunchKernelGGL((k_radix_upsweep<K, B>), dim3(xcd_grid<kXcdBuildRun>(net)),
                       dim3(kSortThreads), 0, st, segs, et, net, tile_hist, shift, src);
    if (sc.nchunks)
      hipLaunchKernelGGL(k_radix_chunk_sum<B>, dim3(grid_for(sc.nchunks)), dim3(kChunkThreads), 0,
                         st, sc.chunks, sc.nchunks, tile_hist, sc.chunk_rows);
    hipLaunchKernelGGL(k_radix_scan<B>, dim3(grid_for(nseg)), dim3(kScanThreads), 0, st, segs,
                       nseg, tile_hist, tile_off, sc.chunk_rows);
    if (sc.nchunks)
      hipLaunchKernelGGL(k_radix_chunk_offsets<B>, dim3(grid_for(sc.nchunks)), dim3(kChunkThreads),
                         0, st, sc.chunks, sc.nchunks, tile_hist, sc.chunk_rows, tile_off);
    hipLaunchKernelGGL((k_radix_downsweep<K, B>), dim3(xcd_grid<kXcdBuildRun>(net)),
                       dim3(kSortThreads), 0, st, segs, et, net, tile_off, shift, src);
    src ^= 1;
  }
}

template <class K>
static void sort_width(hipStream_t st, int bits, const SegDesc* segs, uint32_t nseg,
                       const Tile* et, uint32_t net, const SortScratch& sc) {
  if (bits == 10)
    sort_passes<K, 10>(st, segs, nseg, et, net, sc);
  else if (bits == 9)
    sort_passes<K, 9>(st, segs, nseg, et, net, sc);
  else
    sort_passes<K, 8>(st, segs, nseg, et, net, sc);
}

hipError_t launch_sort(hipStream_t st, const Mode& md, const SegDesc* segs, uint32_t nseg,
                       const Tile* et, uint32_t net, const SortScratch& sc) {
  if (md.radix_bits < 8 || md.radix_bits > kRadixBitsMax) return hipErrorInvalidValue;
  if (md.key63)
    sort_width<uint64_t>(st, md.radix_bits, segs, nseg, et, net, sc);
  else
    sort_width<uint32_t>(st, md.radix_bits, segs, nseg, et, net, sc);
  return hipGetLastError();
}

hipError_t launch_gather(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net) {
  hipLaunchKernelGGL(k_gather, dim3(xcd_grid<kXcdBuildRun>(net)), dim3(kGatherThreads), 0, st,
                     segs, et, net);
  return hipGetLastError();
}

hipError_t launch_refresh_leaves(hipStream_t st, const SegDesc* segs, const Tile* et,
                                 uint32_t net) {
  hipLaunchKernelGGL(k_gather, dim3(xcd_grid<kXcdBuildRun>(net)), dim3(kGatherThreads), 0, st,
                     segs, et, net);
  return hipGetLastError();
}

hipError_t launch_flatten(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net,
                          const FlatOut& out) {
  hipLaunchKernelGGL(k_flatten, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net, out);
  return hipGetLastError();
}

hipError_t launch_karras(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* et,
                         uint32_t net) {
  if (md.key63)
    hipLaunchKernelGGL(k_karras<uint64_t>, dim3(xcd_grid<kXcdBuildRun>(net)), dim3(kKarrasThreads),
                       0, st, segs, et, net);
  else
    hipLaunchKernelGGL(k_karras<uint32_t>, dim3(xcd_grid<kXcdBuildRun>(net)), dim3(kKarrasThreads),
                       0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_aggregate(hipStream_t st, const Mode& md, const SegDesc* segs, uint32_t nseg,
                            const Tile* at, uint32_t nat, const Tile* mt, uint32_t nmt) {
  const dim3 tg(grid_for(nat)), tb(kAggThreads), sg(grid_for(nseg)), sb(kTopThreads);
  const dim3 mg(grid_for(nmt));
  if (md.quad) {
    hipLaunchKernelGGL(k_aggregate_tile<true>, tg, tb, 0, st, segs, at, nat, (int)md.group);
    if (nmt) hipLaunchKernelGGL(k_aggregate_mid<true>, mg, sb, 0, st, segs, mt, nmt);
    hipLaunchKernelGGL(k_aggregate_top<true>, sg, sb, 0, st, segs, nseg);
  } else {
    hipLaunchKernelGGL(k_aggregate_tile<false>, tg, tb, 0, st, segs, at, nat, (int)md.group);
    if (nmt) hipLaunchKernelGGL(k_aggregate_mid<false>, mg, sb, 0, st, segs, mt, nmt);
    hipLaunchKernelGGL(k_aggregate_top<false>, sg, sb, 0, st, segs, nseg);
  }
  return hipGetLastError();
}

template <bool KAHAN, bool COUNT, bool WIDE>
static void walk_group_q(hipStream_t st, bool quad, bool far, const SegDesc* segs, const Tile* gt,
                         uint32_t ngt, int mode) {
  // one block per walk tile (up to 2^25 blocks of one wave): a block that took a second
  // tile a grid's length later would mix two distant regions in its XCD's L2
  const dim3 g(xcd_grid(ngt, 1u << 25)), b(kGroup);
  if (quad && far)
    hipLaunchKernelGGL((k_walk_group<KAHAN, COUNT, true, true, WIDE>), g, b, 0, st, segs, gt, ngt,
                       mode);
  else if (quad)
    hipLaunchKernelGGL((k_walk_group<KAHAN, COUNT, true, false, WIDE>), g, b, 0, st, segs, gt,
                       ngt, mode);
  else
    hipLaunchKernelGGL((k_walk_group<KAHAN, COUNT, false, false, WIDE>), g, b, 0, st, segs, gt,
                       ngt, mode);
}

template <bool KAHAN, bool COUNT>
static void walk_group_w(hipStream_t st, bool wide, bool quad, bool far, const SegDesc* segs,
                         const Tile* gt, uint32_t ngt, int mode) {
  if (wide) walk_group_q<KAHAN, COUNT, true>(st, quad, far, segs, gt, ngt, mode);
  else walk_group_q<KAHAN, COUNT, false>(st, quad, far, segs, gt, ngt, mode);
}

hipError_t launch_walk(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* wt,
                       uint32_t nwt, const Tile* gt, uint32_t ngt, int mode, bool count,
                       bool wide) {
  if (md.group) {
    if (md.kahan) {
      if (count) walk_group_w<true, true>(st, wide, md.quad, md.far, segs, gt, ngt, mode);
      else walk_group_w<true, false>(st, wide, md.quad, md.far, segs, gt, ngt, mode);
    } else {
      if (count) walk_group_w<false, true>(st, wide, md.quad, md.far, segs, gt, ngt, mode);
      else walk_group_w<false, false>(st, wide, md.quad, md.far, segs, gt, ngt, mode);
    }
    return hipGetLastError();
  }
  const dim3 g(xcd_grid<kXcdLaneRun>(nwt)), b(kBlock);
  if (md.kahan) {
    if (count)
      hipLaunchKernelGGL((k_walk_lane<true, true>), g, b, 0, st, segs, wt, nwt, mode);
    else
      hipLaunchKernelGGL((k_walk_lane<true, false>), g, b, 0, st, segs, wt, nwt, mode);
  } else {
    if (count)
      hipLaunchKernelGGL((k_walk_lane<false, true>), g, b, 0, st, segs, wt, nwt, mode);
    else
      hipLaunchKernelGGL((k_walk_lane<false, false>), g, b, 0, st, segs, wt, nwt, mode);
  }
  return hipGetLastError();
}

hipError_t launch_diag(hipStream_t st, const SegDesc* segs, uint32_t nseg, const Tile* et,
                       uint32_t net, double* partial, double* out) {
  hipLaunchKernelGGL(k_diag_partial, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net,
                     partial);
  hipLaunchKernelGGL(k_diag_final, dim3(nseg), dim3(64), 0, st, segs, nseg, partial, out);
  return hipGetLastError();
}

DirectPlan direct_plan(uint32_t n, uint32_t nt, size_t partial_bytes) {
  DirectPlan p{1u, n ? n : 1u};
  if (nt == 0 || n == 0) return p;
  // about 2048 blocks of 256 targets x one source chunk fill the 256 CUs several times
  const uint32_t tblocks = (nt + kDsTile - 1) / kDsTile;
  uint32_t c = (2048u + tblocks - 1) / tblocks;
  const uint32_t by_src = (n + kDsTile - 1) / kDsTile;  // at least one LDS tile per chunk
  if (c > by_src) c = by_src;
  if (c > 2048u) c = 2048u;
  const size_t fit = partial_bytes / (3 * sizeof(double) * (size_t)nt);
  if ((size_t)c > fit) c = (uint32_t)fit;
  if (c < 2u) return p;
  uint32_t len = (n + c - 1) / c;
  len = (len + kDsTile - 1) / kDsTile * kDsTile;
  p.chunks = (n + len - 1) / len;
  p.chunk_len = len;
  if (p.chunks < 2u) p = DirectPlan{1u, n};
  return p;
}

hipError_t launch_direct_f64(hipStream_t st, const double* x, const double* m, uint32_t n,
                             const uint32_t* targets, uint32_t nt, double g, double eps2,
                             const DirectPlan& plan, double* partial, double* out) {
  if (nt == 0) return hipSuccess;
  const uint32_t tblocks = (nt + kDsTile - 1) / kDsTile;
  if (plan.chunks < 2u) {
    hipLaunchKernelGGL(k_direct_f64, dim3(tblocks, 1), dim3(kDsTile), 0, st, x, m, n, targets,
                       nt, g, eps2, plan.chunk_len, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_direct_f64, dim3(tblocks, plan.chunks), dim3(kDsTile), 0, st, x, m, n,
                     targets, nt, g, eps2, plan.chunk_len, partial);
  const uint32_t rblocks = (uint32_t)((3 * (size_t)nt + kDsTile - 1) / kDsTile);
  hipLaunchKernelGGL(k_direct_reduce, dim3(rblocks), dim3(kDsTile), 0, st, partial, nt,
                     plan.chunks, out);
  return hipGetLastError();
}

hipError_t launch_potential_f64(hipStream_t st, const double* x, const double* m, uint32_t n,
                                double g, double eps2, double* partial, double* out) {
  const uint32_t nb = (n + kDsTile - 1) / kDsTile;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_potential_f64, dim3(nb), dim3(kDsTile), 0, st, x, m, n, g, eps2, partial);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1), 0, st, partial, nb, out);
  return hipGetLastError();
}

}  // namespace glx
