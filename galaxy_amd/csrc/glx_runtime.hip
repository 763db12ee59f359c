// Host runtime: device segments, batches and launch sequences.
#include "glx_runtime.h"

#include <mutex>

#include "../../include/galaxy_lbvh.h"

namespace glx {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    fail(e == hipErrorOutOfMemory ? GLX_ENOMEM : GLX_EDEVICE,
         std::string(what) + ": " + hipGetErrorString(e));
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
  size_t x, v, m, acc, acc_out, k0, k1, v0, v1, leaf, node, left, right, pleaf, pint, rs, hi,
      flag, bbox, bounds, counters, total;
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
  L.k0 = c.take<uint32_t>(N);
  L.k1 = c.take<uint32_t>(N);
  L.v0 = c.take<uint32_t>(N);
  L.v1 = c.take<uint32_t>(N);
  L.leaf = c.take<float4>(N);
  L.node = c.take<Node>(N);
  L.left = c.take<uint32_t>(N);
  L.right = c.take<uint32_t>(N);
  L.pleaf = c.take<uint32_t>(N);
  L.pint = c.take<uint32_t>(N);
  L.rs = c.take<uint32_t>(N);
  L.hi = c.take<uint32_t>(N);
  L.flag = c.take<uint32_t>(N);
  L.bbox = c.take<unsigned long long>(6);
  L.bounds = c.take<double>(8);
  L.counters = c.take<unsigned long long>(4);
  L.total = c.off;
  return L;
}
}  // namespace

size_t Segment::bytes_for(uint32_t n, bool with_acc_out) {
  return layout_for(n, with_acc_out).total;
}

Segment::~Segment() {
  if (block_) (void)hipFree(block_);
}

void Segment::reserve(uint32_t n, bool with_acc_out) {
  if (block_ && n <= capacity && (has_acc_out || !with_acc_out)) return;
  uint32_t cap = n;
  if (block_ && n > capacity) cap = n > capacity + capacity / 2 ? n : capacity + capacity / 2;
  const bool acc_out = with_acc_out || has_acc_out;
  const Layout L = layout_for(cap, acc_out);
  void* blk = nullptr;
  check(hipMalloc(&blk, L.total), "hipMalloc(segment)");
  // flags must start at zero (the bottom-up pass keeps them zero afterwards)
  hipError_t e = hipMemset(static_cast<char*>(blk) + L.flag, 0, sizeof(uint32_t) * (cap ? cap : 1));
  if (e == hipSuccess) e = hipMemset(static_cast<char*>(blk) + L.counters, 0, 32);
  if (e != hipSuccess) {
    (void)hipFree(blk);
    check(e, "hipMemset(segment)");
  }
  if (block_) (void)hipFree(block_);
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
  d.acc_out = nullptr;
  d.keys[0] = reinterpret_cast<uint32_t*>(b + L.k0);
  d.keys[1] = reinterpret_cast<uint32_t*>(b + L.k1);
  d.vals[0] = reinterpret_cast<uint32_t*>(b + L.v0);
  d.vals[1] = reinterpret_cast<uint32_t*>(b + L.v1);
  d.leaf = reinterpret_cast<float4*>(b + L.leaf);
  d.node = reinterpret_cast<Node*>(b + L.node);
  d.left = reinterpret_cast<uint32_t*>(b + L.left);
  d.right = reinterpret_cast<uint32_t*>(b + L.right);
  d.parent_leaf = reinterpret_cast<uint32_t*>(b + L.pleaf);
  d.parent_int = reinterpret_cast<uint32_t*>(b + L.pint);
  d.rs = reinterpret_cast<uint32_t*>(b + L.rs);
  d.hi = reinterpret_cast<uint32_t*>(b + L.hi);
  d.flag = reinterpret_cast<uint32_t*>(b + L.flag);
  d.bbox = reinterpret_cast<unsigned long long*>(b + L.bbox);
  d.bounds = reinterpret_cast<double*>(b + L.bounds);
  d.counters = reinterpret_cast<unsigned long long*>(b + L.counters);
  d.acc_out = acc_out ? reinterpret_cast<double*>(b + L.acc_out) : nullptr;
  desc = d;
}

// ----------------------------------------------------------------------------------
// Batch
// ----------------------------------------------------------------------------------

Batch::~Batch() {
  (void)hipFree(segs_);
  (void)hipFree(etiles_);
  (void)hipFree(wtiles_);
  (void)hipFree(tile_hist_);
  (void)hipFree(tile_off_);
  (void)hipFree(diag_partial_);
  (void)hipFree(diag_out_);
}

template <class T>
static void grow(T*& p, size_t& cap, size_t need) {
  if (need <= cap && p) return;
  size_t c = need > cap * 2 ? need : cap * 2;
  if (c == 0) c = 1;
  T* q = nullptr;
  check(hipMalloc(&q, c * sizeof(T)), "hipMalloc(batch table)");
  (void)hipFree(p);
  p = q;
  cap = c;
}

void Batch::set_segments(const std::vector<SegDesc>& descs) {
  host_ = descs;
  nseg_ = (uint32_t)descs.size();
  std::vector<Tile> et, wt;
  for (uint32_t s = 0; s < nseg_; ++s) {
    SegDesc& d = host_[s];
    d.sort_tile0 = (uint32_t)et.size();
    for (uint32_t st = 0; st < d.n; st += kSortTile)
      et.push_back(Tile{s, st, (d.n - st) < (uint32_t)kSortTile ? d.n - st : (uint32_t)kSortTile, 0});
    d.sort_tiles = (uint32_t)et.size() - d.sort_tile0;
    for (uint32_t st = 0; st < d.n; st += kBlock)
      wt.push_back(Tile{s, st, (d.n - st) < (uint32_t)kBlock ? d.n - st : (uint32_t)kBlock, 0});
  }
  n_et_ = (uint32_t)et.size();
  n_wt_ = (uint32_t)wt.size();
  size_t cs = cap_segs_;
  grow(segs_, cs, nseg_ ? nseg_ : 1);
  if (cs != cap_segs_) {
    (void)hipFree(diag_out_);
    diag_out_ = nullptr;
    check(hipMalloc(&diag_out_, cs * kDiagRecord * sizeof(double)), "hipMalloc(diag)");
  }
  cap_segs_ = cs;
  size_t ce = cap_et_;
  grow(etiles_, ce, n_et_ ? n_et_ : 1);
  if (ce != cap_et_ || !tile_hist_) {
    (void)hipFree(tile_hist_);
    (void)hipFree(tile_off_);
    (void)hipFree(diag_partial_);
    tile_hist_ = tile_off_ = nullptr;
    diag_partial_ = nullptr;
    check(hipMalloc(&tile_hist_, ce * kRadixBins * sizeof(uint32_t)), "hipMalloc(hist)");
    check(hipMalloc(&tile_off_, ce * kRadixBins * sizeof(uint32_t)), "hipMalloc(hist)");
    check(hipMalloc(&diag_partial_, ce * kDiagRecord * sizeof(double)), "hipMalloc(diag)");
  }
  cap_et_ = ce;
  grow(wtiles_, cap_wt_, n_wt_ ? n_wt_ : 1);
  if (nseg_)
    check(hipMemcpyAsync(segs_, host_.data(), nseg_ * sizeof(SegDesc), hipMemcpyHostToDevice,
                         stream_), "upload segments");
  if (n_et_)
    check(hipMemcpyAsync(etiles_, et.data(), n_et_ * sizeof(Tile), hipMemcpyHostToDevice, stream_),
          "upload tiles");
  if (n_wt_)
    check(hipMemcpyAsync(wtiles_, wt.data(), n_wt_ * sizeof(Tile), hipMemcpyHostToDevice, stream_),
          "upload tiles");
  check(hipStreamSynchronize(stream_), "set_segments");
}

void Batch::reset_bbox() { check(launch_seg_reset(stream_, segs_, nseg_), "seg_reset"); }
void Batch::bbox() { check(launch_bbox(stream_, segs_, etiles_, n_et_), "bbox"); }
void Batch::kick_drift_bbox() {
  check(launch_kick_drift_bbox(stream_, segs_, etiles_, n_et_), "kick_drift");
}
void Batch::encode(uint32_t* lanes_out) {
  check(launch_encode(stream_, segs_, etiles_, n_et_, lanes_out), "encode");
}
void Batch::bounds_only() { check(launch_bounds_only(stream_, segs_, nseg_), "bounds"); }
void Batch::iota() { check(launch_iota(stream_, segs_, etiles_, n_et_), "iota"); }
void Batch::sort() {
  check(launch_sort(stream_, segs_, nseg_, etiles_, n_et_, tile_hist_, tile_off_), "sort");
}
void Batch::gather() { check(launch_gather(stream_, segs_, etiles_, n_et_), "gather"); }
void Batch::karras() { check(launch_karras(stream_, segs_, etiles_, n_et_), "karras"); }
void Batch::aggregate() { check(launch_aggregate(stream_, segs_, etiles_, n_et_), "aggregate"); }
void Batch::walk(int mode, bool kahan, bool count) {
  check(launch_walk(stream_, segs_, wtiles_, n_wt_, mode, kahan, count), "walk");
}

void Batch::build_and_walk(int mode, bool kahan, bool count) {
  encode();
  sort();
  gather();
  karras();
  aggregate();
  walk(mode, kahan, count);
}

void Batch::diag(double* host_out) {
  check(launch_diag(stream_, segs_, nseg_, etiles_, n_et_, diag_partial_, diag_out_), "diag");
  check(hipMemcpyAsync(host_out, diag_out_, nseg_ * kDiagRecord * sizeof(double),
                       hipMemcpyDeviceToHost, stream_), "diag download");
  check(hipStreamSynchronize(stream_), "diag");
}

}  // namespace glx
