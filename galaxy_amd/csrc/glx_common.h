// Shared device/host definitions of the MI355X LBVH gravity step.
//
// Memory model: every system ("segment") owns its own device arrays, listed in a
// SegDesc that lives in device memory; a batched launch covers many segments by
// iterating over a table of Tile descriptors {segment, first element, count}. A
// single system is the one-segment case of the same kernels.
//
// Per-segment index spaces (all local to the segment):
//   particles        i in [0, n)          original order (the caller's index order)
//   sorted leaves    k in [0, n)          Morton order; leaf k holds particle vals[0][k]
//   internal nodes   j in [0, n-1)        Karras numbering; unified index n + j
//   unified index    u < n  -> leaf u,    u >= n -> internal node u - n, root = n
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace glx {

constexpr uint32_t kNoParent = 0xFFFFFFFFu;  // solvers/src/lbvh.rs:244
constexpr uint32_t kEnd = 0xFFFFFFFFu;       // walk terminator
constexpr int kRadixBits = 8;                          // default sort digit width
constexpr int kRadixBitsMax = 10;                      // widest digit (8, 9 or 10 bits)
constexpr int kRadixBinsMax = 1 << kRadixBitsMax;
constexpr int kSortThreads = 512;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per element tile
constexpr uint32_t kScanChunk = 32;     // sort tiles per chunk of the multi-block digit scan
constexpr uint32_t kScanChunkMin = 256; // segments of more sort tiles than this scan in chunks
constexpr int kBlock = 256;                            // threads per block of element kernels
constexpr int kAggTile = 1024;                         // leaves per bottom-up tile (LDS pass)
constexpr uint32_t kMidTiles = 32;                     // tiles per bottom-up super-tile
constexpr uint32_t kMidSpan = kMidTiles * kAggTile;    // leaves per super-tile (middle pass)
#ifndef GLX_MID_MIN
#define GLX_MID_MIN (8u * kMidSpan)
#endif
// Systems of more than kMidMin bodies run the middle bottom-up pass (below it the top
// pass alone is faster than two dependent climbing kernels).
constexpr uint32_t kMidMin = GLX_MID_MIN;
constexpr int kGroup = 64;                             // targets per group walk (one wave)

// Walk record of one internal node, split by use into two arrays of 32-B records: the
// classifier record, which every node test of the group walk loads, and the box record,
// which only the tests that need the node's box load (the per-target walk's inside test,
// the group walk's disjoint-box test at theta > 1 and its bucket test). Siblings are
// adjacent (children g and g + 1), so one 128-B line holds the classifier records of two
// sibling pairs.
struct __align__(16) Node {
  float4 a;  // centre of mass (origin-relative f32), mass
  float4 h;  // squared acceptance radius (s/theta + delta)^2, children as walk entries
             // (enc_entries, as bits); then, with quadrupoles, the group walk's squared
             // monopole-band radius and the node's radius about its com, otherwise the
             // next node after this subtree (unified or kEnd)
};
struct __align__(16) NodeBox {
  float4 lo;  // AABB min (origin-relative f32), first leaf (as bits)
  float4 hi;  // AABB max (origin-relative f32), last leaf (as bits)
};

// Children of an internal node in one word: the split position g (left child g, right
// child g + 1, each a leaf or an internal node) and two leaf flags (n <= 2^29).
constexpr uint32_t kLeftLeaf = 1u << 30, kRightLeaf = 1u << 31, kSplitMask = kLeftLeaf - 1u;
__host__ __device__ inline uint32_t enc_children(uint32_t cl, uint32_t cr, uint32_t n) {
  const uint32_t g = cl < n ? cl : cl - n;
  return g | (cl < n ? kLeftLeaf : 0u) | (cr < n ? kRightLeaf : 0u);
}
__host__ __device__ inline uint32_t left_child(uint32_t e, uint32_t n) {
  return (e & kLeftLeaf) ? (e & kSplitMask) : n + (e & kSplitMask);
}
__host__ __device__ inline uint32_t right_child(uint32_t e, uint32_t n) {
  return (e & kRightLeaf) ? (e & kSplitMask) + 1u : n + (e & kSplitMask) + 1u;
}

// The group walk addresses a segment's leaves, classifier records and quadrupoles by
// 32-bit byte offsets from its leaf array when all three lie within 4 GiB of it (the
// segment layout keeps them adjacent, 80 B per body: up to about 5e7 bodies); other
// segments take the walk's 64-bit-address form.
__host__ __device__ inline bool walk_off32(uint64_t leaf, uint64_t node, uint64_t quad, uint32_t n) {
  const uint64_t lim = 0xFFFFFFFFull, dn = node - leaf, dq = quad - leaf;
  return dn <= lim && dq <= lim && 16ull * n <= lim && dn + 32ull * n <= lim &&
         dq + 32ull * n <= lim;
}

// Children in the node record as walk entries (leaf g -> g, internal node g -> g | 2^31):
// the left child's entry, with bit 30 set when the right child (g + 1 for a split at g) is
// of the other kind (n <= 2^29). c is a unified index (leaf c < n, internal node c - n).
constexpr uint32_t kEntryNode = 1u << 31, kEntryFlip = 1u << 30;
__host__ __device__ inline uint32_t walk_entry(uint32_t c, uint32_t n) {
  return c < n ? c : (c - n) | kEntryNode;
}
__host__ __device__ inline uint32_t enc_entries(uint32_t cl, uint32_t cr, uint32_t n) {
  return walk_entry(cl, n) | ((cl < n) != (cr < n) ? kEntryFlip : 0u);
}
__host__ __device__ inline uint32_t left_entry(uint32_t w) { return w & ~kEntryFlip; }
__host__ __device__ inline uint32_t right_entry(uint32_t w) {
  return (w & ~kEntryFlip) + 1u + ((w & kEntryFlip) << 1);  // the flip toggles the kind bit
}

// Build/walk configuration shared by all segments of a launch sequence.
struct Mode {
  bool key63 = false;    // 63-bit Morton keys (2^21 cells per axis) instead of 30-bit
  bool group = false;    // wave-cooperative group walk instead of the per-target walk
  bool quad = false;     // quadrupole moments (group walk only)
  bool kahan = false;    // Kahan-compensated accumulation
  bool far = false;      // per-group local expansion of well-separated nodes (group walk)
  int radix_bits = kRadixBits;  // sort digit width: 8, 9 or 10 bits per pass
};

// Passes of the LSD sort: key bits over the digit width. An odd count starts the sort in
// buffer 1 so that the sorted keys and payload always end in buffer 0.
__host__ __device__ constexpr int sort_passes_for(bool key63, int radix_bits) {
  return ((key63 ? 63 : 30) + radix_bits - 1) / radix_bits;
}
__host__ __device__ constexpr int sort_start(bool key63, int radix_bits) {
  return sort_passes_for(key63, radix_bits) & 1;
}

// walk accumulation and drift precision: kPrecF32 and kPrecKahan walk in f32 coordinates
// relative to the root-cube centre and drift positions in f64; kPrecAbs32 is the plain
// all-f32 path (f32 coordinates relative to the coordinate origin, f32 drift)
enum : uint32_t { kPrecF32 = 0, kPrecKahan = 1, kPrecAbs32 = 2 };

// counters[] slots
enum : int {
  kCntInteractions = 0,  // point-mass terms (leaf bodies and monopole nodes), per target
  kCntVisits = 1,        // node tests
  kCntRoots = 2,         // (unused)
  kCntOverflow = 3,      // group walks that fell back to the per-target walk
  kCntQuad = 4,          // quadrupole-node terms, per target
  kCntMaxStack = 5,      // deepest group-walk stack, entries
  kCntFar = 6,           // nodes folded into a group's local expansion, per group
  kCntRounds = 7,        // group-walk rounds, summed over groups
  kCounters = 8
};

struct SegDesc {
  // resident state (original particle order)
  double* x;        // 3n positions, f64, interleaved xyz
  float* v;         // 3n velocities, f32
  double* m;        // n masses, f64
  float* acc;       // 3n accelerations (f32, already x G) cached between steps
  double* acc_out;  // optional 3n f64 output of a one-shot force evaluation
  // build scratch
  void* keys[2];         // sort keys: uint32_t (30-bit) or uint64_t (63-bit)
  uint32_t* vals[2];     // sort payload (original index)
  float4* leaf;          // n sorted leaves: origin-relative f32 position, mass
  Node* node;            // n-1 internal-node classifier records
  NodeBox* box;          // n-1 internal-node box records
  float4* quad;          // 2(n-1) quadrupole moments {xx,yy,zz,xy},{xz,yz,-,-} or null
  uint4* topo;           // n-1 internal node j: first leaf, last leaf of its range,
                         // children (split | leaf flags, enc_children), parent (unified;
                         // kNoParent for the root) — one load for the bottom-up climbs
  uint32_t* parent_leaf; // n parent of leaf k (unified)
  uint32_t* rs;          // n-1 right child of the internal node whose split is s
  uint32_t* flag;        // n-1 arrival counters of the top bottom-up pass (kept at 0)
  uint32_t* roots;       // n maximal subtrees of each bottom-up tile, per tile at its first
                         // leaf (rcount[tile] of them), handed to the middle or top pass
  uint32_t* roots2;      // n the same per super-tile (rcount2 of them), for the top pass
  uint32_t* rcount;      // ceil(n / kAggTile) roots handed over per tile
  uint32_t* rcount2;     // ceil(n / kMidSpan) roots handed over per super-tile
  uint32_t* gend;        // n walk groups: last leaf of the group starting at k, else kEnd
  const uint8_t* active; // optional walk mask over the original indices (null: walk all)
  unsigned long long* bbox;      // 6 order-preserving encodings: min xyz, max xyz
  double* bounds;                // 8: bmin xyz, size, scale, origin xyz
  unsigned long long* counters;  // kCounters: see kCnt*
  uint32_t n;
  uint32_t sort_tile0, sort_tiles;  // element tiles (kSortTile) of this segment
  uint32_t scan_chunk0, scan_chunks;  // its chunks of kScanChunk tiles (0: scanned whole)
  uint32_t precision;
  uint32_t leaf_size;  // group walk: nodes of at most this many bodies are summed body by body
  uint32_t group_size;  // group walk: bodies per walk group (subtrees of at most this many)
  uint32_t group_lanes; // group walk: most lanes per body in groups of at most 32 bodies
  uint32_t stack_cap;   // group walk: stack entries before the per-target fallback (0: all)
  float g, eps2, theta, half_dt;
  float far2;  // squared far-field ratio: (node radius + group radius)^2 <= far2 * R^2
  float mono_inv;  // 1 / (theta mono_ratio): accepted nodes this far are monopoles (0: off)
  double dt;
};

struct Tile {
  uint32_t seg, start, count, pad;
};

}  // namespace glx
