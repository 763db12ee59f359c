// Host runtime: device segments, batches and launch sequences.
#include "glx_runtime.h"

#include <cstdlib>
#include <mutex>

#include "galaxy_lbvh.h"

namespace glx {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    fail(e == hipErrorOutOfMemory ? GLX_ENOMEM : GLX_EDEVICE,
         std::string(what) + ": " + hipGetErrorString(e));
}

std::recursive_mutex& device_memory_mutex() {
  static std::recursive_mutex mu;
  return mu;
}

void dev_malloc(void** p, size_t bytes, const char* what) {
  std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());
  *p = nullptr;
  check(hipMalloc(p, bytes), what);
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());
  (void)hipFree(p);
}

void ensure_device() {
  static std::once_flag once;
  static int status = 0;
  static std::string msg;
  std::call_once(once, [] {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
      status = GLX_EDEVICE;
      msg = "no HIP device available";
      return;
    }
    e = hipSetDevice(0);
    if (e != hipSuccess) {
      status = GLX_EDEVICE;
      msg = std::string("hipSetDevice: ") + hipGetErrorString(e);
    }
  });
  if (status) fail(status, msg);
}

static inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ----------------------------------------------------------------------------------
// Segment
// ----------------------------------------------------------------------------------

namespace {
struct Carve {
  size_t off = 0;
  template <class T>
  size_t take(size_t count) {
    const size_t o = off;
    off = align_up(off + count * sizeof(T), 256);
    return o;
  }
};

struct Layout {
  size_t x, v, m, acc, acc_out, k0, k1, v0, v1, leaf, node, box, quad, topo, pleaf, rs, flag,
      roots, roots2, rcount, rcount2, gend, bbox, bounds, counters, total;
};

Layout layout_for(uint32_t n, bool with_acc_out) {
  const size_t N = n ? n : 1;
  Carve c;
  Layout L;
  L.x = c.take<double>(3 * N);
  L.v = c.take<float>(3 * N);
  L.m = c.take<double>(N);
  L.acc = c.take<float>(3 * N);
  L.acc_out = with_acc_out ? c.take<double>(3 * N) : 0;
  L.k0 = c.take<uint64_t>(N);
  L.k1 = c.take<uint64_t>(N);
  L.v0 = c.take<uint32_t>(N);
  L.v1 = c.take<uint32_t>(N);
  // leaves, classifier records and quadrupoles adjacent: the group walk addresses all
  // three with 32-bit byte offsets from the leaf array (80 B per body)
  L.leaf = c.take<float4>(N);
  L.node = c.take<Node>(N);
  L.quad = c.take<float4>(2 * N);
  L.box = c.take<NodeBox>(N);
  L.topo = c.take<uint4>(N);
  L.pleaf = c.take<uint32_t>(N);
  L.rs = c.take<uint32_t>(N);
  L.flag = c.take<uint32_t>(N);
  L.roots = c.take<uint32_t>(N);
  L.roots2 = c.take<uint32_t>(N);
  L.rcount = c.take<uint32_t>((N + kAggTile - 1) / kAggTile);
  L.rcount2 = c.take<uint32_t>((N + kMidSpan - 1) / kMidSpan);
  L.gend = c.take<uint32_t>(N);
  L.bbox = c.take<unsigned long long>(6);
  L.bounds = c.take<double>(8);
  L.counters = c.take<unsigned long long>(kCounters);
  L.total = c.off;
  return L;
}
}  // namespace

size_t Segment::bytes_for(uint32_t n, bool with_acc_out) {
  return layout_for(n, with_acc_out).total;
}

Segment::~Segment() { dev_free(block_); }

void Segment::reserve(uint32_t n, bool with_acc_out) {
  if (block_ && n <= capacity && (has_acc_out || !with_acc_out)) return;
  uint32_t cap = n;
  if (block_ && n > capacity) cap = n > capacity + capacity / 2 ? n : capacity + capacity / 2;
  const bool acc_out = with_acc_out || has_acc_out;
  const Layout L = layout_for(cap, acc_out);
  std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());
  void* blk = nullptr;
  dev_malloc(&blk, L.total, "hipMalloc(segment)");
  // arrival flags and counters must start at zero (kernels keep them zero afterwards)
  hipError_t e = hipMemset(static_cast<char*>(blk) + L.flag, 0, sizeof(uint32_t) * (cap ? cap : 1));
  if (e == hipSuccess)
    e = hipMemset(static_cast<char*>(blk) + L.counters, 0, kCounters * sizeof(unsigned long long));
  if (e != hipSuccess) {
    dev_free(blk);
    check(e, "hipMemset(segment)");
  }
  dev_free(block_);
  block_ = blk;
  bytes_ = L.total;
  capacity = cap;
  has_acc_out = acc_out;
  char* b = static_cast<char*>(blk);
  SegDesc d = desc;
  d.x = reinterpret_cast<double*>(b + L.x);
  d.v = reinterpret_cast<float*>(b + L.v);
  d.m = reinterpret_cast<double*>(b + L.m);
  d.acc = reinterpret_cast<float*>(b + L.acc);
  d.acc_out = acc_out ? reinterpret_cast<double*>(b + L.acc_out) : nullptr;
  d.keys[0] = b + L.k0;
  d.keys[1] = b + L.k1;
  d.vals[0] = reinterpret_cast<uint32_t*>(b + L.v0);
  d.vals[1] = reinterpret_cast<uint32_t*>(b + L.v1);
  d.leaf = reinterpret_cast<float4*>(b + L.leaf);
  d.node = reinterpret_cast<Node*>(b + L.node);
  d.box = reinterpret_cast<NodeBox*>(b + L.box);
  d.quad = reinterpret_cast<float4*>(b + L.quad);
  d.topo = reinterpret_cast<uint4*>(b + L.topo);
  d.parent_leaf = reinterpret_cast<uint32_t*>(b + L.pleaf);
  d.rs = reinterpret_cast<uint32_t*>(b + L.rs);
  d.flag = reinterpret_cast<uint32_t*>(b + L.flag);
  d.roots = reinterpret_cast<uint32_t*>(b + L.roots);
  d.roots2 = reinterpret_cast<uint32_t*>(b + L.roots2);
  d.rcount = reinterpret_cast<uint32_t*>(b + L.rcount);
  d.rcount2 = reinterpret_cast<uint32_t*>(b + L.rcount2);
  d.gend = reinterpret_cast<uint32_t*>(b + L.gend);
  d.bbox = reinterpret_cast<unsigned long long*>(b + L.bbox);
  d.bounds = reinterpret_cast<double*>(b + L.bounds);
  d.counters = reinterpret_cast<unsigned long long*>(b + L.counters);
  desc = d;
}

// ----------------------------------------------------------------------------------
// Batch
// ----------------------------------------------------------------------------------

// The side stream and its events exist from construction, so that a step graph captured
// on the batch stream never creates them mid-capture.
Batch::Batch(hipStream_t st) : stream_(st) {
  check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreate");
  if (hipEventCreateWithFlags(&fork_, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&join_, hipEventDisableTiming) != hipSuccess) {
    if (fork_) (void)hipEventDestroy(fork_);
    (void)hipStreamDestroy(side_);
    fail(GLX_EDEVICE, "hipEventCreate failed");
  }
}

Batch::~Batch() {
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
  if (side_) (void)hipStreamDestroy(side_);
  dev_free(segs_);
  dev_free(etiles_);
  dev_free(atiles_);
  dev_free(mtiles_);
  dev_free(ctiles_);
  dev_free(chunk_rows_);
  dev_free(wtiles_);
  dev_free(gtiles_);
  dev_free(tile_hist_);
  dev_free(tile_off_);
  dev_free(diag_partial_);
  dev_free(diag_out_);
}

template <class T>
static void grow(T*& p, size_t& cap, size_t need) {
  if (need <= cap && p) return;
  size_t c = need > cap * 2 ? need : cap * 2;
  if (c == 0) c = 1;
  T* q = nullptr;
  dev_malloc(reinterpret_cast<void**>(&q), c * sizeof(T), "hipMalloc(batch table)");
  dev_free(p);
  p = q;
  cap = c;
}

static void push_tiles(std::vector<Tile>& v, uint32_t seg, uint32_t n, uint32_t size) {
  for (uint32_t st = 0; st < n; st += size)
    v.push_back(Tile{seg, st, (n - st) < size ? n - st : size, 0});
}

void Batch::set_segments(const std::vector<SegDesc>& descs) {
  host_ = descs;
  nseg_ = (uint32_t)descs.size();
  std::vector<Tile> et, at, mt, ct, wt, gt;
  // GLX_WALK_WIDE=1 takes the 64-bit-address walk for every segment (the two forms give
  // the same results bit for bit; the tests compare them on systems of any size)
  const char* force_wide = std::getenv("GLX_WALK_WIDE");
  wide_ = force_wide && force_wide[0] == '1';
  // GLX_WALK_STACK_CAP=k caps the group walk's stack at k entries, so that the tests can
  // drive groups into the per-target fallback (default: the full LDS stack)
  const char* cap = std::getenv("GLX_WALK_STACK_CAP");
  const uint32_t stack_cap = cap ? (uint32_t)std::strtoul(cap, nullptr, 10) : 0u;
  for (uint32_t s = 0; s < nseg_; ++s) {
    SegDesc& d = host_[s];
    d.stack_cap = stack_cap;
    if (!walk_off32((uint64_t)d.leaf, (uint64_t)d.node, (uint64_t)d.quad, d.n)) wide_ = true;
    d.sort_tile0 = (uint32_t)et.size();
    push_tiles(et, s, d.n, kSortTile);
    d.sort_tiles = (uint32_t)et.size() - d.sort_tile0;
    d.scan_chunk0 = (uint32_t)ct.size();
    if (d.sort_tiles > kScanChunkMin)  // chunks of kScanChunk of the segment's sort tiles
      for (uint32_t c = 0; c < d.sort_tiles; c += kScanChunk)
        ct.push_back(Tile{s, d.sort_tile0 + c, d.sort_tiles - c < kScanChunk ? d.sort_tiles - c : kScanChunk, 0});
    d.scan_chunks = (uint32_t)ct.size() - d.scan_chunk0;
    push_tiles(at, s, d.n, kAggTile);
    if (d.n > kMidMin) push_tiles(mt, s, d.n, kMidSpan);
    push_tiles(wt, s, d.n, kBlock);
    push_tiles(gt, s, d.n, kGroup);
  }
  n_et_ = (uint32_t)et.size();
  n_at_ = (uint32_t)at.size();
  n_mt_ = (uint32_t)mt.size();
  n_ct_ = (uint32_t)ct.size();
  n_wt_ = (uint32_t)wt.size();
  n_gt_ = (uint32_t)gt.size();
  size_t cs = cap_segs_;
  grow(segs_, cs, nseg_ ? nseg_ : 1);
  if (cs != cap_segs_ || !diag_out_) {
    dev_free(diag_out_);
    diag_out_ = nullptr;
    dev_malloc(reinterpret_cast<void**>(&diag_out_), cs * kDiagRecord * sizeof(double),
               "hipMalloc(diag)");
  }
  cap_segs_ = cs;
  size_t ce = cap_et_;
  grow(etiles_, ce, n_et_ ? n_et_ : 1);
  if (ce != cap_et_ || !tile_hist_) {
    dev_free(tile_hist_);
    dev_free(tile_off_);
    dev_free(diag_partial_);
    tile_hist_ = tile_off_ = nullptr;
    diag_partial_ = nullptr;
    dev_malloc(reinterpret_cast<void**>(&tile_hist_), ce * kRadixBinsMax * sizeof(uint32_t),
               "hipMalloc(hist)");
    dev_malloc(reinterpret_cast<void**>(&tile_off_), ce * kRadixBinsMax * sizeof(uint32_t),
               "hipMalloc(hist)");
    dev_malloc(reinterpret_cast<void**>(&diag_partial_), ce * kDiagRecord * sizeof(double),
               "hipMalloc(diag)");
  }
  cap_et_ = ce;
  grow(atiles_, cap_at_, n_at_ ? n_at_ : 1);
  grow(mtiles_, cap_mt_, n_mt_ ? n_mt_ : 1);
  size_t cc = cap_ct_;
  grow(ctiles_, cc, n_ct_ ? n_ct_ : 1);
  if (cc != cap_ct_ || !chunk_rows_) {
    dev_free(chunk_rows_);
    chunk_rows_ = nullptr;
    dev_malloc(reinterpret_cast<void**>(&chunk_rows_), cc * kRadixBinsMax * sizeof(uint32_t),
               "hipMalloc(scan chunks)");
  }
  cap_ct_ = cc;
  grow(wtiles_, cap_wt_, n_wt_ ? n_wt_ : 1);
  grow(gtiles_, cap_gt_, n_gt_ ? n_gt_ : 1);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes)
      check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "upload table");
  };
  up(segs_, host_.data(), nseg_ * sizeof(SegDesc));
  up(etiles_, et.data(), n_et_ * sizeof(Tile));
  up(atiles_, at.data(), n_at_ * sizeof(Tile));
  up(mtiles_, mt.data(), n_mt_ * sizeof(Tile));
  up(ctiles_, ct.data(), n_ct_ * sizeof(Tile));
  up(wtiles_, wt.data(), n_wt_ * sizeof(Tile));
  up(gtiles_, gt.data(), n_gt_ * sizeof(Tile));
  check(hipStreamSynchronize(stream_), "set_segments");
}

void Batch::set_dt(double dt) {
  for (auto& d : host_) {
    d.dt = dt;
    d.half_dt = (float)(0.5 * dt);
  }
  if (nseg_)
    check(hipMemcpyAsync(segs_, host_.data(), nseg_ * sizeof(SegDesc), hipMemcpyHostToDevice,
                         stream_),
          "upload segments");
  check(hipStreamSynchronize(stream_), "set_dt");
}

void Batch::set_active(const uint8_t* mask) {
  for (auto& d : host_) d.active = mask;
  if (nseg_)
    check(hipMemcpyAsync(segs_, host_.data(), nseg_ * sizeof(SegDesc), hipMemcpyHostToDevice,
                         stream_),
          "upload segments");
  check(hipStreamSynchronize(stream_), "set_active");
}

void Batch::reset_bbox() { check(launch_seg_reset(stream_, segs_, nseg_), "seg_reset"); }
void Batch::bbox() { check(launch_bbox(stream_, segs_, etiles_, n_et_), "bbox"); }
void Batch::kick_drift_bbox() {
  check(launch_kick_drift_bbox(stream_, segs_, etiles_, n_et_), "kick_drift");
}
void Batch::encode(uint32_t* lanes_out) {
  check(launch_encode(stream_, mode_, segs_, etiles_, n_et_, lanes_out), "encode");
}
void Batch::bounds_only() { check(launch_bounds_only(stream_, mode_, segs_, nseg_), "bounds"); }
void Batch::iota() { check(launch_iota(stream_, mode_, segs_, etiles_, n_et_), "iota"); }
void Batch::sort() {
  const SortScratch sc{tile_hist_, tile_off_, ctiles_, n_ct_, chunk_rows_};
  check(launch_sort(stream_, mode_, segs_, nseg_, etiles_, n_et_, sc), "sort");
}
void Batch::gather() { check(launch_gather(stream_, segs_, etiles_, n_et_), "gather"); }
void Batch::refresh_leaves() {
  check(launch_refresh_leaves(stream_, segs_, etiles_, n_et_), "refresh leaves");
}
void Batch::karras() { check(launch_karras(stream_, mode_, segs_, etiles_, n_et_), "karras"); }
void Batch::aggregate() {
  check(launch_aggregate(stream_, mode_, segs_, nseg_, atiles_, n_at_, mtiles_, n_mt_), "aggregate");
}
void Batch::flatten(const FlatOut& out) {
  check(launch_flatten(stream_, segs_, etiles_, n_et_, out), "flatten");
}
void Batch::walk(int mode, bool count) {
  check(launch_walk(stream_, mode_, segs_, wtiles_, n_wt_, gtiles_, n_gt_, mode, count, wide_),
        "walk");
}

void Batch::build_and_walk(int mode, bool count) {
  encode();
  sort();
  // the gather (HBM-bound) beside the tree build (latency-bound): both read only the
  // sort's output and write disjoint arrays; the bottom-up pass needs both
  check(hipEventRecord(fork_, stream_), "hipEventRecord");
  check(hipStreamWaitEvent(side_, fork_, 0), "hipStreamWaitEvent");
  check(launch_gather(side_, segs_, etiles_, n_et_), "gather");
  check(hipEventRecord(join_, side_), "hipEventRecord");
  karras();
  check(hipStreamWaitEvent(stream_, join_, 0), "hipStreamWaitEvent");
  aggregate();
  walk(mode, count);
}

void Batch::diag(double* host_out) {
  check(launch_diag(stream_, segs_, nseg_, etiles_, n_et_, diag_partial_, diag_out_), "diag");
  check(hipMemcpyAsync(host_out, diag_out_, nseg_ * kDiagRecord * sizeof(double),
                       hipMemcpyDeviceToHost, stream_), "diag download");
  check(hipStreamSynchronize(stream_), "diag");
}

}  // namespace glx
