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
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide exclusive scan of one u32 per thread (blockDim == kBlock == 256).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh_wave,
                                                         uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) sh_wave[w] = inc;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int k = 0; k < kBlock / 64; ++k) {
    uint32_t c = sh_wave[k];
    if (k < w) pre += c;
    tot += c;
  }
  __syncthreads();
  if (total) *total = tot;
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
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return expand10(x) | (expand10(y) << 1) | (expand10(z) << 2);
}

// solvers/src/lbvh.rs:110-120: the padded root cube of the segment's bbox, in f64
// with the reference's operation order (no contraction). origin = cube centre.
struct Bounds {
  double bmin[3], size, scale, origin[3];
};
__device__ Bounds seg_bounds(const unsigned long long* bbox) {
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
  b.scale = (double)(1u << kMortonBits) / b.size;
  return b;
}

// solvers/src/lbvh.rs:56-64: floor, clamp below at 0 (NaN -> 0), saturate at 1023.
__device__ __forceinline__ uint32_t quant_lane(double v) {
  const double f = fmax(floor(v), 0.0);
  return f >= 1023.0 ? 1023u : (uint32_t)f;
}

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

// Opening half-kick with a(x_n), f64 drift, and the bbox of the drifted positions
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
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const size_t i = tl.start + e;
      for (int a = 0; a < 3; ++a) {
        const float v = s.v[3 * i + a] + s.acc[3 * i + a] * h;
        s.v[3 * i + a] = v;
        const double x = s.x[3 * i + a] + (double)v * dt;
        s.x[3 * i + a] = x;
        lo[a] = fmin(lo[a], x);
        hi[a] = fmax(hi[a], x);
      }
    }
    block_commit_bbox(lo, hi, s.bbox, sh);
  }
}

// ----------------------------------------------------------------------------------
// 2. Morton encode (bit-exact with solvers/src/lbvh.rs:105-129)
// ----------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_encode(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
         uint32_t* __restrict__ lanes_out) {
#pragma clang fp contract(off)
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const Bounds b = seg_bounds(s.bbox);
    if (tl.start == 0 && threadIdx.x == 0) {
      s.bounds[0] = b.bmin[0];
      s.bounds[1] = b.bmin[1];
      s.bounds[2] = b.bmin[2];
      s.bounds[3] = b.size;
      s.bounds[4] = b.scale;
      s.bounds[5] = b.origin[0];
      s.bounds[6] = b.origin[1];
      s.bounds[7] = b.origin[2];
    }
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const uint32_t i = tl.start + e;
      uint32_t l[3];
      for (int a = 0; a < 3; ++a) l[a] = quant_lane((s.x[3 * (size_t)i + a] - b.bmin[a]) * b.scale);
      if (lanes_out) {
        lanes_out[3 * (size_t)i] = l[0];
        lanes_out[3 * (size_t)i + 1] = l[1];
        lanes_out[3 * (size_t)i + 2] = l[2];
      }
      s.keys[0][i] = morton3(l[0], l[1], l[2]);
      s.vals[0][i] = i;
    }
  }
}

// Bounds only (used when keys are supplied by the caller).
__global__ void k_bounds(const SegDesc* __restrict__ segs, uint32_t nseg) {
  const uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nseg) return;
  const SegDesc& s = segs[sidx];
  const Bounds b = seg_bounds(s.bbox);
  for (int a = 0; a < 3; ++a) {
    s.bounds[a] = b.bmin[a];
    s.bounds[5 + a] = b.origin[a];
  }
  s.bounds[3] = b.size;
  s.bounds[4] = b.scale;
}

__global__ void __launch_bounds__(kBlock)
k_iota(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) s.vals[0][tl.start + e] = tl.start + e;
  }
}

// ----------------------------------------------------------------------------------
// 3. stable LSD radix sort, segmented reduce-then-scan (per pass: upsweep, scan,
//    downsweep). Pure integer; reproduces the (code, index) order of
//    solvers/src/lbvh.rs:143-147 because the payload starts as iota and every pass
//    is stable.
// ----------------------------------------------------------------------------------

__global__ void __launch_bounds__(kSortThreads)
k_radix_upsweep(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
                uint32_t* __restrict__ tile_hist, int shift, int src) {
  __shared__ uint32_t hist[kRadixBins];
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const uint32_t* keys = s.keys[src] + tl.start;
    hist[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < tl.count; e += kSortThreads)
      atomicAdd(&hist[(keys[e] >> shift) & (kRadixBins - 1)], 1u);
    __syncthreads();
    tile_hist[(size_t)t * kRadixBins + threadIdx.x] = hist[threadIdx.x];
    __syncthreads();
  }
}

// One block per segment: per digit, exclusive scan over the segment's tiles plus the
// segment-wide digit base (exclusive scan over digits of the digit totals).
__global__ void __launch_bounds__(kSortThreads)
k_radix_scan(const SegDesc* __restrict__ segs, uint32_t nseg, const uint32_t* __restrict__ tile_hist,
             uint32_t* __restrict__ tile_off) {
  __shared__ uint32_t shw[kSortThreads / 64];
  const uint32_t d = threadIdx.x;
  for (uint32_t sidx = blockIdx.x; sidx < nseg; sidx += gridDim.x) {
    const SegDesc& s = segs[sidx];
    const uint32_t t0 = s.sort_tile0, t1 = s.sort_tile0 + s.sort_tiles;
    uint32_t tot = 0;
    for (uint32_t t = t0; t < t1; ++t) tot += tile_hist[(size_t)t * kRadixBins + d];
    uint32_t run = block_exclusive_scan(tot, shw, nullptr);
    for (uint32_t t = t0; t < t1; ++t) {
      const uint32_t h = tile_hist[(size_t)t * kRadixBins + d];
      tile_off[(size_t)t * kRadixBins + d] = run;
      run += h;
    }
  }
}

__global__ void __launch_bounds__(kSortThreads)
k_radix_downsweep(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
                  const uint32_t* __restrict__ tile_off, int shift, int src) {
  constexpr int W = kSortThreads / 64;
  constexpr int PER_WAVE = kSortTile / W;
  __shared__ uint32_t wcnt[W][kRadixBins];
  __shared__ uint32_t dstart[kRadixBins];
  __shared__ uint32_t skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  __shared__ uint32_t shw[W];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const uint32_t* keys_in = s.keys[src] + tl.start;
    const uint32_t* vals_in = s.vals[src] + tl.start;
    uint32_t* keys_out = s.keys[src ^ 1];
    uint32_t* vals_out = s.vals[src ^ 1];
    for (int k = 0; k < W; ++k) wcnt[k][threadIdx.x] = 0;
    __syncthreads();

    uint32_t key[kSortItems], val[kSortItems], dig[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      const uint32_t e = w * PER_WAVE + j * 64 + lane;  // striped: index order = (j, lane)
      const bool valid = e < tl.count;
      key[j] = valid ? keys_in[e] : 0u;
      val[j] = valid ? vals_in[e] : 0u;
      const uint32_t dd = valid ? ((key[j] >> shift) & (kRadixBins - 1)) : (uint32_t)kRadixBins;
      // lanes holding the same digit (bit kRadixBits separates the invalid tail)
      unsigned long long peers = ~0ull;
#pragma unroll
      for (int b = 0; b <= kRadixBits; ++b) {
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
      const uint32_t d = threadIdx.x;
      uint32_t c[W], tot = 0;
      for (int k = 0; k < W; ++k) {
        c[k] = wcnt[k][d];
        tot += c[k];
      }
      uint32_t run = 0;
      for (int k = 0; k < W; ++k) {
        wcnt[k][d] = run;
        run += c[k];
      }
      dstart[d] = block_exclusive_scan(tot, shw, nullptr);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      if (dig[j] < (uint32_t)kRadixBins) {
        const uint32_t p = dstart[dig[j]] + wcnt[w][dig[j]] + rank[j];
        skey[p] = key[j];
        sval[p] = val[j];
      }
    }
    __syncthreads();
    const uint32_t* off = tile_off + (size_t)t * kRadixBins;
    for (uint32_t p = threadIdx.x; p < tl.count; p += kSortThreads) {
      const uint32_t k = skey[p];
      const uint32_t d = (k >> shift) & (kRadixBins - 1);
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

__global__ void __launch_bounds__(kBlock)
k_gather(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const double o0 = s.bounds[5], o1 = s.bounds[6], o2 = s.bounds[7];
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const uint32_t k = tl.start + e;
      const size_t i = min(s.vals[0][k], s.n - 1);
      float4 lf;
      lf.x = (float)(s.x[3 * i] - o0);
      lf.y = (float)(s.x[3 * i + 1] - o1);
      lf.z = (float)(s.x[3 * i + 2] - o2);
      lf.w = (float)s.m[i];
      s.leaf[k] = lf;
      if (s.n == 1) s.parent_leaf[0] = kNoParent;
    }
  }
}

// ----------------------------------------------------------------------------------
// 5. Karras radix-tree build (solvers/src/lbvh.rs:153-240,276-306), signed int
//    indices; each internal node also records its range end and publishes itself
//    as the right child of its split position, from which the skip pointers follow.
// ----------------------------------------------------------------------------------

__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ codes, int n, int a, int b) {
  if (b < 0 || b >= n) return -1;
  const uint32_t ca = codes[a], cb = codes[b];
  if (ca == cb) return 32 + __clz((int)((uint32_t)a ^ (uint32_t)b));
  return __clz((int)(ca ^ cb));
}

__global__ void __launch_bounds__(kBlock)
k_karras(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const int n = (int)s.n;
    const uint32_t* codes = s.keys[0];
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const int i = (int)(tl.start + e);
      if (i >= n - 1) continue;
      const int dir = kdelta(codes, n, i, i + 1) > kdelta(codes, n, i, i - 1) ? 1 : -1;
      const int dmin = kdelta(codes, n, i, i - dir);
      int lmax = 2;
      while (kdelta(codes, n, i, i + lmax * dir) > dmin) lmax *= 2;
      int l = 0;
      for (int tt = lmax / 2; tt > 0; tt /= 2)
        if (kdelta(codes, n, i, i + (l + tt) * dir) > dmin) l += tt;
      const int j = i + l * dir;
      const int dnode = kdelta(codes, n, i, j);
      int sp = 0, tt = l;
      for (;;) {
        tt = (tt + 1) / 2;
        if (kdelta(codes, n, i, i + (sp + tt) * dir) > dnode) sp += tt;
        if (tt <= 1) break;
      }
      const int gamma = i + sp * dir + (dir < 0 ? dir : 0);
      const int lo = i < j ? i : j, hi = i < j ? j : i;
      const uint32_t lu = (lo == gamma) ? (uint32_t)gamma : (uint32_t)(n + gamma);
      const uint32_t ru = (hi == gamma + 1) ? (uint32_t)(gamma + 1) : (uint32_t)(n + gamma + 1);
      const uint32_t me = (uint32_t)(n + i);
      s.left[i] = lu;
      s.right[i] = ru;
      if (lu < (uint32_t)n) s.parent_leaf[lu] = me; else s.parent_int[lu - n] = me;
      if (ru < (uint32_t)n) s.parent_leaf[ru] = me; else s.parent_int[ru - n] = me;
      s.rs[gamma] = ru;
      s.hi[i] = (uint32_t)hi;
      if (i == 0) s.parent_int[0] = kNoParent;
    }
  }
}

// ----------------------------------------------------------------------------------
// 6. bottom-up aggregate with arrival flags (fold order fixed: left, right;
//    solvers/src/lbvh.rs:623-636). One thread per leaf climbs while it is the second
//    arrival at a parent; the flag is reset by that second arrival so the array is
//    all-zero again for the next build.
// ----------------------------------------------------------------------------------

struct Agg {
  float3 com, mn, mx;
  float mass;
};

__device__ __forceinline__ Agg load_child(const SegDesc& s, uint32_t c, uint32_t n) {
  Agg r;
  if (c < n) {
    const float4 p = s.leaf[c];
    r.com = make_float3(p.x, p.y, p.z);
    r.mn = r.com;
    r.mx = r.com;
    r.mass = p.w;
  } else {
    const Node* nd = &s.node[c - n];
    const float4 a = nd->a, b = nd->b, cc = nd->c;
    r.com = make_float3(a.x, a.y, a.z);
    r.mass = a.w;
    r.mn = make_float3(b.x, b.y, b.z);
    r.mx = make_float3(cc.x, cc.y, cc.z);
  }
  return r;
}

__global__ void __launch_bounds__(kBlock)
k_aggregate(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    const uint32_t n = s.n;
    if (n < 2) continue;
    const float inv_theta = 1.0f / s.theta;
    for (uint32_t e = threadIdx.x; e < tl.count; e += kBlock) {
      const uint32_t k = tl.start + e;
      uint32_t p = s.parent_leaf[k];
      // a climb visits each ancestor once; the bound only guards corrupted input
      for (uint32_t guard = 0; p != kNoParent && guard < n; ++guard) {
        const uint32_t j = p - n;
        if (j >= n - 1) break;
        __threadfence();
        const uint32_t old =
            __hip_atomic_fetch_add(&s.flag[j], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0u) break;  // first arrival: the sibling's thread finishes this node
        __threadfence();
        __hip_atomic_store(&s.flag[j], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t lc = s.left[j], rc = s.right[j];
        const Agg la = load_child(s, lc, n), ra = load_child(s, rc, n);
        Agg r;
        r.mass = la.mass + ra.mass;
        if (r.mass > 0.0f) {
          r.com.x = (la.com.x * la.mass + ra.com.x * ra.mass) / r.mass;
          r.com.y = (la.com.y * la.mass + ra.com.y * ra.mass) / r.mass;
          r.com.z = (la.com.z * la.mass + ra.com.z * ra.mass) / r.mass;
        } else {
          r.com.x = (la.mn.x + la.mx.x + ra.mn.x + ra.mx.x) * 0.25f;
          r.com.y = (la.mn.y + la.mx.y + ra.mn.y + ra.mx.y) * 0.25f;
          r.com.z = (la.mn.z + la.mx.z + ra.mn.z + ra.mx.z) * 0.25f;
        }
        r.mn = make_float3(fminf(la.mn.x, ra.mn.x), fminf(la.mn.y, ra.mn.y), fminf(la.mn.z, ra.mn.z));
        r.mx = make_float3(fmaxf(la.mx.x, ra.mx.x), fmaxf(la.mx.y, ra.mx.y), fmaxf(la.mx.z, ra.mx.z));
        // opening geometry (solvers/src/lbvh.rs:602-611, 511-514): accept iff
        // theta*(d - delta) >= s  <=>  d >= s/theta + delta
        const float cx = (r.mn.x + r.mx.x) * 0.5f, cy = (r.mn.y + r.mx.y) * 0.5f,
                    cz = (r.mn.z + r.mx.z) * 0.5f;
        const float hx = (r.mx.x - r.mn.x) * 0.5f, hy = (r.mx.y - r.mn.y) * 0.5f,
                    hz = (r.mx.z - r.mn.z) * 0.5f;
        const float size = 2.0f * fmaxf(hx, fmaxf(hy, hz));
        const float ddx = r.com.x - cx, ddy = r.com.y - cy, ddz = r.com.z - cz;
        const float delta = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        const float rcrit = size * inv_theta + delta;
        const uint32_t h = s.hi[j];
        Node nd;
        nd.a = make_float4(r.com.x, r.com.y, r.com.z, r.mass);
        nd.b = make_float4(r.mn.x, r.mn.y, r.mn.z, rcrit * rcrit);
        nd.c = make_float4(r.mx.x, r.mx.y, r.mx.z, __uint_as_float(lc));
        nd.d = make_uint4(h == n - 1 ? kEnd : s.rs[h], 0u, 0u, 0u);
        s.node[j] = nd;
        p = s.parent_int[j];
      }
    }
  }
}

// ----------------------------------------------------------------------------------
// 7. walk (solvers/src/lbvh.rs:469-524): stackless skip-pointer traversal in DFS
//    pre-order (left child first; after a subtree, its skip pointer), one target per
//    lane in Morton order, Plummer-softened monopole, f32 arithmetic in
//    origin-relative coordinates; optional Kahan-compensated accumulation.
//    Epilogue: scatter a (x G) to the original index; resident mode also applies the
//    closing half-kick (core/src/integrator.rs:70-73).
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

template <bool KAHAN, bool COUNT>
__global__ void __launch_bounds__(kBlock)
k_walk(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
       int mode) {
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    if (tl.start == 0 && threadIdx.x < 6) bbox_neutral(s.bbox, threadIdx.x);
    if (threadIdx.x >= tl.count) continue;
    const uint32_t n = s.n;
    const uint32_t k = tl.start + threadIdx.x;
    const float4 me = s.leaf[k];
    const float eps2 = s.eps2;
    const float4* __restrict__ leaf = s.leaf;
    const Node* __restrict__ node = s.node;
    const uint32_t* __restrict__ rs = s.rs;
    Acc3 a = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    unsigned long long n_int = 0, n_vis = 0;
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
        const float4 A = nd->a;
        const uint32_t nxt = nd->d.x;
        if (!(A.w > 0.0f)) {
          u = nxt;
          continue;
        }
        const float4 B = nd->b, Cc = nd->c;
        const bool inside = (B.x <= me.x) & (me.x <= Cc.x) & (B.y <= me.y) & (me.y <= Cc.y) &
                            (B.z <= me.z) & (me.z <= Cc.z);
        const float dx = A.x - me.x, dy = A.y - me.y, dz = A.z - me.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (!inside && d2 >= B.w) {
          const float r2 = d2 + eps2;
          const float ir = rsqrtf(r2);
          const float w = A.w * ir * ir * ir;
          acc_add<KAHAN>(a, dx * w, dy * w, dz * w);
          if (COUNT) ++n_int;
          u = nxt;
        } else {
          u = __float_as_uint(Cc.w);
        }
      }
    }
    const float g = s.g;
    const float ax = (a.x + (KAHAN ? -a.cx : 0.f)) * g;
    const float ay = (a.y + (KAHAN ? -a.cy : 0.f)) * g;
    const float az = (a.z + (KAHAN ? -a.cz : 0.f)) * g;
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
    if (COUNT) {
      atomicAdd(&s.counters[0], n_int);
      atomicAdd(&s.counters[1], n_vis);
    }
  }
}

// ----------------------------------------------------------------------------------
// 8. FP64 diagnostics (core/src/diagnostics.rs:8-68): per-tile partial sums in a
//    fixed reduction order, then a fixed-order sum over each segment's tiles.
//    Partial record: KE, P xyz, L xyz, sum m*x xyz, sum m  (11 doubles).
// ----------------------------------------------------------------------------------

constexpr int kDiagWords = 11;

__global__ void __launch_bounds__(kBlock)
k_diag_partial(const SegDesc* __restrict__ segs, const Tile* __restrict__ tiles, uint32_t ntiles,
               double* __restrict__ partial) {
  __shared__ double sh[kDiagWords][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const SegDesc& s = segs[tl.seg];
    double acc[kDiagWords];
    for (int q = 0; q < kDiagWords; ++q) acc[q] = 0.0;
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
    for (int q = 0; q < kDiagWords; ++q) {
      const double r = wave_sum(acc[q]);
      if (lane == 0) sh[q][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < kDiagWords) {
      double r = 0.0;
      for (int k = 0; k < kBlock / 64; ++k) r += sh[threadIdx.x][k];
      partial[(size_t)t * kDiagWords + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

__global__ void k_diag_final(const SegDesc* __restrict__ segs, uint32_t nseg,
                             const double* __restrict__ partial, double* __restrict__ out) {
  const uint32_t sidx = blockIdx.x;
  if (sidx >= nseg || threadIdx.x >= kDiagWords) return;
  const SegDesc& s = segs[sidx];
  double r = 0.0;
  for (uint32_t t = s.sort_tile0; t < s.sort_tile0 + s.sort_tiles; ++t)
    r += partial[(size_t)t * kDiagWords + threadIdx.x];
  out[(size_t)sidx * kDiagWords + threadIdx.x] = r;
}

// ----------------------------------------------------------------------------------
// 9. FP64 direct summation (solvers/src/direct_sum.rs:25-45 pair kernel) onto a
//    target subset, and the FP64 softened potential (solvers/src/potential.rs:17-27)
//    as per-block partials. Sources are staged through LDS.
// ----------------------------------------------------------------------------------

constexpr int kDsTile = 256;

__global__ void __launch_bounds__(kDsTile)
k_direct_f64(const double* __restrict__ x, const double* __restrict__ m, uint32_t n,
             const uint32_t* __restrict__ targets, uint32_t nt, double g, double eps2,
             double* __restrict__ out) {
  __shared__ double sx[kDsTile], sy[kDsTile], sz[kDsTile], sm[kDsTile];
  const uint32_t t = blockIdx.x * kDsTile + threadIdx.x;
  const bool act = t < nt;
  const uint32_t i = act ? targets[t] : 0u;
  const double xi = act ? x[3 * (size_t)i] : 0.0, yi = act ? x[3 * (size_t)i + 1] : 0.0,
               zi = act ? x[3 * (size_t)i + 2] : 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (uint32_t base = 0; base < n; base += kDsTile) {
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
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double inv_r3 = 1.0 / (r2 * sqrt(r2));
      const double w = (base + q == i) ? 0.0 : g * inv_r3 * sm[q];
      ax += dx * w;
      ay += dy * w;
      az += dz * w;
    }
    __syncthreads();
  }
  if (act) {
    out[3 * (size_t)t] = ax;
    out[3 * (size_t)t + 1] = ay;
    out[3 * (size_t)t + 2] = az;
  }
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
      if (base + q <= i) continue;
      const double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
      const double r = sqrt(dx * dx + dy * dy + dz * dz + eps2);
      u -= g * mi * sm[q] / r;
    }
    __syncthreads();
  }
  if (!act) u = 0.0;
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

// ----------------------------------------------------------------------------------
// host-side launch wrappers
// ----------------------------------------------------------------------------------

static inline uint32_t grid_for(uint32_t ntiles, uint32_t cap = 65535u) {
  return ntiles == 0 ? 1u : (ntiles < cap ? ntiles : cap);
}

__global__ void k_seg_reset(const SegDesc* __restrict__ segs, uint32_t nseg) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nseg * 6u) return;
  bbox_neutral(segs[q / 6u].bbox, (int)(q % 6u));
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

hipError_t launch_encode(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net,
                         uint32_t* lanes_out) {
  hipLaunchKernelGGL(k_encode, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net, lanes_out);
  return hipGetLastError();
}

hipError_t launch_bounds_only(hipStream_t st, const SegDesc* segs, uint32_t nseg) {
  hipLaunchKernelGGL(k_bounds, dim3((nseg + 63) / 64), dim3(64), 0, st, segs, nseg);
  return hipGetLastError();
}

hipError_t launch_iota(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net) {
  hipLaunchKernelGGL(k_iota, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_sort(hipStream_t st, const SegDesc* segs, uint32_t nseg, const Tile* et,
                       uint32_t net, uint32_t* tile_hist, uint32_t* tile_off) {
  int src = 0;
  for (int pass = 0; pass < kRadixPasses; ++pass) {
    const int shift = pass * kRadixBits;
    hipLaunchKernelGGL(k_radix_upsweep, dim3(grid_for(net)), dim3(kSortThreads), 0, st, segs, et,
                       net, tile_hist, shift, src);
    hipLaunchKernelGGL(k_radix_scan, dim3(grid_for(nseg)), dim3(kSortThreads), 0, st, segs, nseg,
                       tile_hist, tile_off);
    hipLaunchKernelGGL(k_radix_downsweep, dim3(grid_for(net)), dim3(kSortThreads), 0, st, segs, et,
                       net, tile_off, shift, src);
    src ^= 1;
  }
  return hipGetLastError();
}

hipError_t launch_gather(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net) {
  hipLaunchKernelGGL(k_gather, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_karras(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net) {
  hipLaunchKernelGGL(k_karras, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_aggregate(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net) {
  hipLaunchKernelGGL(k_aggregate, dim3(grid_for(net)), dim3(kBlock), 0, st, segs, et, net);
  return hipGetLastError();
}

hipError_t launch_walk(hipStream_t st, const SegDesc* segs, const Tile* wt, uint32_t nwt,
                       int mode, bool kahan, bool count) {
  const dim3 g(grid_for(nwt)), b(kBlock);
  if (kahan) {
    if (count)
      hipLaunchKernelGGL((k_walk<true, true>), g, b, 0, st, segs, wt, nwt, mode);
    else
      hipLaunchKernelGGL((k_walk<true, false>), g, b, 0, st, segs, wt, nwt, mode);
  } else {
    if (count)
      hipLaunchKernelGGL((k_walk<false, true>), g, b, 0, st, segs, wt, nwt, mode);
    else
      hipLaunchKernelGGL((k_walk<false, false>), g, b, 0, st, segs, wt, nwt, mode);
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

hipError_t launch_direct_f64(hipStream_t st, const double* x, const double* m, uint32_t n,
                             const uint32_t* targets, uint32_t nt, double g, double eps2,
                             double* out) {
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(k_direct_f64, dim3((nt + kDsTile - 1) / kDsTile), dim3(kDsTile), 0, st, x, m,
                     n, targets, nt, g, eps2, out);
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
