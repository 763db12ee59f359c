// Host runtime of the LBVH gravity step: device segments, batches of segments with
// their tile tables, and the launch sequences of one force evaluation / KDK step.
#pragma once
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "glx_kernels.h"

namespace glx {

// Failure inside the runtime: carries a GLX_* status and a message. Never crosses
// the C ABI (every entry point catches it).
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void check(hipError_t e, const char* what);
void ensure_device();

// Device allocation and graph capture exclude each other process-wide: a synchronous
// runtime call (allocation, free, a null-stream memset) made in one thread while another
// thread captures a step graph would join the capture or invalidate it. Every hipMalloc,
// hipFree and null-stream memset of the library, and every capture, holds this lock.
std::recursive_mutex& device_memory_mutex();
void dev_malloc(void** p, size_t bytes, const char* what);  // throws GLX_ENOMEM / EDEVICE
void dev_free(void* p);                                       // null is a no-op

// One system's device arrays (a single allocation carved into the arrays of
// SegDesc). Capacity grows on demand and never shrinks.
class Segment {
 public:
  Segment() = default;
  ~Segment();
  Segment(const Segment&) = delete;
  Segment& operator=(const Segment&) = delete;

  // Ensure room for n particles; with_acc_out adds the f64 force output buffer.
  void reserve(uint32_t n, bool with_acc_out);
  static size_t bytes_for(uint32_t n, bool with_acc_out);
  size_t bytes() const { return bytes_; }

  SegDesc desc{};  // host copy of the descriptor (pointers into the allocation)
  uint32_t capacity = 0;
  bool has_acc_out = false;

 private:
  void* block_ = nullptr;
  size_t bytes_ = 0;
};

// A set of segments launched together, with the tile tables and the scratch shared
// by one launch sequence.
class Batch {
 public:
  explicit Batch(hipStream_t st);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // Rebuild tables for the given descriptors (copied to the device).
  void set_segments(const std::vector<SegDesc>& descs);
  // Patch per-segment step parameters (dt) on the device without rebuilding tables.
  void set_dt(double dt);
  // Patch the walk mask of every segment (null: walk all).
  void set_active(const uint8_t* mask);
  void set_mode(const Mode& m) { mode_ = m; }
  const Mode& mode() const { return mode_; }
  uint32_t nseg() const { return nseg_; }
  const std::vector<SegDesc>& host_descs() const { return host_; }
  SegDesc* dev_descs() const { return segs_; }

  // Launch sequences (all on the batch stream, no host synchronization).
  void reset_bbox();
  void bbox();
  void kick_drift_bbox();
  void encode(uint32_t* lanes_out = nullptr);
  void bounds_only();
  void iota();
  void sort();
  void gather();
  void refresh_leaves();  // new positions into the cached leaf order (no rebuild)
  void karras();
  void aggregate();
  void flatten(const FlatOut& out);  // pre-order export of the built tree (parity only)
  void walk(int mode, bool count);
  // bbox-independent tail: encode .. walk (bbox must already be accumulated); the gather
  // runs on a side stream beside the tree build (both need only the sort)
  void build_and_walk(int mode, bool count);
  void diag(double* host_out /* nseg * kDiagRecord */);

  hipStream_t stream() const { return stream_; }

 private:
  hipStream_t stream_;
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  Mode mode_{};
  uint32_t nseg_ = 0;
  std::vector<SegDesc> host_;
  SegDesc* segs_ = nullptr;
  Tile* etiles_ = nullptr;  // kSortTile elements per tile
  Tile* atiles_ = nullptr;  // kAggTile leaves per tile (bottom-up pass)
  Tile* mtiles_ = nullptr;  // kMidSpan leaves per super-tile (systems larger than one)
  Tile* ctiles_ = nullptr;  // kScanChunk sort tiles per chunk (segments scanned in chunks)
  uint32_t* chunk_rows_ = nullptr;  // per chunk: digit sums, then digit offsets
  Tile* wtiles_ = nullptr;  // kBlock targets per tile (per-target walk)
  Tile* gtiles_ = nullptr;  // kGroup targets per tile (group walk)
  uint32_t n_et_ = 0, n_at_ = 0, n_mt_ = 0, n_ct_ = 0, n_wt_ = 0, n_gt_ = 0;
  bool wide_ = false;  // a segment needs the group walk's 64-bit-address form
  uint32_t* tile_hist_ = nullptr;
  uint32_t* tile_off_ = nullptr;
  double* diag_partial_ = nullptr;
  double* diag_out_ = nullptr;
  size_t cap_segs_ = 0, cap_et_ = 0, cap_at_ = 0, cap_mt_ = 0, cap_ct_ = 0, cap_wt_ = 0,
         cap_gt_ = 0;
};

}  // namespace glx
