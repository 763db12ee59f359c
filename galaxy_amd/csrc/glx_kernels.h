// Launch wrappers of the device kernels (internal to the library).
#pragma once
#include "glx_common.h"

namespace glx {

constexpr int kDiagRecord = 11;  // KE, P xyz, L xyz, sum m*x xyz, sum m

// Output of the pre-order flatten (2n-1 slots; float arrays of 3 hold xyz per slot).
struct FlatOut {
  uint32_t *next, *body_start, *body_count;
  float *center, *half, *com, *mass, *delta;
};

hipError_t launch_seg_reset(hipStream_t st, const SegDesc* segs, uint32_t nseg);
hipError_t launch_bbox(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net);
hipError_t launch_kick_drift_bbox(hipStream_t st, const SegDesc* segs, const Tile* et,
                                  uint32_t net);
hipError_t launch_encode(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* et,
                         uint32_t net, uint32_t* lanes_out);
hipError_t launch_bounds_only(hipStream_t st, const Mode& md, const SegDesc* segs, uint32_t nseg);
hipError_t launch_iota(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* et,
                       uint32_t net);
// Per-batch scratch of the sort: per-tile digit histograms and offsets (2^B per tile), and
// the chunk table and per-chunk rows of the segments scanned in chunks (nchunks may be 0).
struct SortScratch {
  uint32_t* tile_hist;
  uint32_t* tile_off;
  const Tile* chunks;
  uint32_t nchunks;
  uint32_t* chunk_rows;
};
hipError_t launch_sort(hipStream_t st, const Mode& md, const SegDesc* segs, uint32_t nseg,
                       const Tile* et, uint32_t net, const SortScratch& sc);
hipError_t launch_gather(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net);
hipError_t launch_refresh_leaves(hipStream_t st, const SegDesc* segs, const Tile* et,
                                 uint32_t net);
hipError_t launch_flatten(hipStream_t st, const SegDesc* segs, const Tile* et, uint32_t net,
                          const FlatOut& out);
hipError_t launch_karras(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* et,
                         uint32_t net);
// Bottom-up pass: at = tiles of kAggTile leaves; mt = super-tiles of kMidSpan leaves of the
// systems larger than one super-tile (none: nmt = 0).
hipError_t launch_aggregate(hipStream_t st, const Mode& md, const SegDesc* segs, uint32_t nseg,
                            const Tile* at, uint32_t nat, const Tile* mt, uint32_t nmt);
// wide: some segment's leaves, node records or quadrupoles lie beyond 4 GiB of its leaf
// array (walk_off32 false), so the group walk takes its 64-bit-address form.
hipError_t launch_walk(hipStream_t st, const Mode& md, const SegDesc* segs, const Tile* wt,
                       uint32_t nwt, const Tile* gt, uint32_t ngt, int mode, bool count,
                       bool wide);
hipError_t launch_diag(hipStream_t st, const SegDesc* segs, uint32_t nseg, const Tile* et,
                       uint32_t net, double* partial, double* out);
// FP64 direct sum onto nt targets: the sources split into `chunks` ranges of chunk_len
// (a multiple of the 256-source LDS tile); with two or more chunks the per-chunk sums go
// to partial (3 * nt * chunks doubles) and are added in chunk order.
struct DirectPlan {
  uint32_t chunks, chunk_len;
};
constexpr size_t kDirectPartialBytes = (size_t)16 << 20;  // partial sums of one call, at most
DirectPlan direct_plan(uint32_t n, uint32_t nt, size_t partial_bytes);
hipError_t launch_direct_f64(hipStream_t st, const double* x, const double* m, uint32_t n,
                             const uint32_t* targets, uint32_t nt, double g, double eps2,
                             const DirectPlan& plan, double* partial, double* out);
hipError_t launch_potential_f64(hipStream_t st, const double* x, const double* m, uint32_t n,
                                double g, double eps2, double* partial, double* out);

}  // namespace glx
