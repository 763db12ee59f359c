// Helpers shared by the C-ABI translation units (glx_api.hip, glx_pool.hip): error
// capture, argument validation, parameter mapping and small RAII device objects.
#pragma once
#include <math.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "galaxy_lbvh.h"
#include "glx_runtime.h"

namespace glx {
namespace abi {

// Message of the last failed call on this thread (defined in glx_api.hip).
extern thread_local std::string last_error;

// Run f, converting runtime failures into a GLX_* status (no exception crosses the ABI).
template <class F>
int guard(F&& f) {
  try {
    f();
    last_error.clear();
    return GLX_OK;
  } catch (const Error& e) {
    last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    last_error = "host allocation failed";
    return GLX_ENOMEM;
  } catch (...) {
    last_error = "unexpected internal error";
    return GLX_EDEVICE;
  }
}

constexpr size_t kMaxN = (size_t)1 << 29;
constexpr uint32_t kMaxLeafSize = 32;

inline void validate_params(const glx_params* p) {
  if (!p) fail(GLX_EINVAL, "params is NULL");
  if (!isfinite(p->g)) fail(GLX_EINVAL, "g must be finite");
  if (!isfinite(p->softening) || p->softening < 0.0) fail(GLX_EINVAL, "softening must be >= 0");
  if (!isfinite(p->theta) || !(p->theta > 0.0)) fail(GLX_EINVAL, "theta must be > 0");
  if (p->leaf_size > kMaxLeafSize) fail(GLX_EINVAL, "leaf_size must be at most 32");
  if (p->leaf_size > 1 && p->walk != GLX_WALK_GROUP)
    fail(GLX_EINVAL, "leaf_size > 1 needs the group walk");
  if (p->multipole != 0 && p->multipole != 2) fail(GLX_EINVAL, "multipole must be 0 or 2");
  if (p->precision > GLX_PREC_ABS32) fail(GLX_EINVAL, "unknown precision mode");
  if (p->morton_bits != 0 && p->morton_bits != 30 && p->morton_bits != 63)
    fail(GLX_EINVAL, "morton_bits must be 30 or 63");
  if (p->walk > GLX_WALK_GROUP) fail(GLX_EINVAL, "unknown walk");
  if (p->multipole == 2 && p->walk != GLX_WALK_GROUP)
    fail(GLX_EINVAL, "quadrupole moments need the group walk");
  if (!isfinite(p->far_theta) || p->far_theta < 0.0 || p->far_theta >= 1.0)
    fail(GLX_EINVAL, "far_theta must be in [0, 1)");
  if (p->far_theta > 0.0 && (p->walk != GLX_WALK_GROUP || p->multipole != 2))
    fail(GLX_EINVAL, "far_theta needs the group walk with quadrupoles");
  if (!isfinite(p->mono_ratio) || p->mono_ratio < 0.0 || p->mono_ratio >= 1.0)
    fail(GLX_EINVAL, "mono_ratio must be in [0, 1)");
  if (p->mono_ratio > 0.0 && p->multipole != 2)
    fail(GLX_EINVAL, "mono_ratio needs quadrupoles");
  if (p->radix_bits != 0 && (p->radix_bits < 8 || p->radix_bits > (uint32_t)kRadixBitsMax))
    fail(GLX_EINVAL, "radix_bits must be 8, 9 or 10");
  if (p->group_size != 0 && (p->group_size < 8 || p->group_size > (uint32_t)kGroup))
    fail(GLX_EINVAL, "group_size must be in [8, 64]");
  if (p->group_lanes > 4) fail(GLX_EINVAL, "group_lanes must be in [1, 4]");
}

inline Mode mode_from(const glx_params& p) {
  Mode m;
  m.key63 = p.morton_bits == 63;
  m.group = p.walk == GLX_WALK_GROUP;
  m.quad = p.multipole == 2;
  m.kahan = p.precision == GLX_PREC_KAHAN;
  m.far = p.far_theta > 0.0;
  m.radix_bits = p.radix_bits ? (int)p.radix_bits : kRadixBits;
  return m;
}

inline void validate_n(size_t n) {
  if (n > kMaxN) fail(GLX_EINVAL, "n exceeds 2^29 particles per system");
}

// Arguments of the FP64 direct sum: targets index the n sources.
inline void validate_direct(size_t n, const double* pos, const double* mass,
                            const uint32_t* targets, size_t n_targets, double* acc_out) {
  validate_n(n);
  if (n_targets == 0) return;
  if (!pos || !mass || !targets || !acc_out) fail(GLX_EINVAL, "NULL array");
  if (n_targets > 0xFFFFFFFFu) fail(GLX_EINVAL, "too many targets");
  for (size_t t = 0; t < n_targets; ++t)
    if (targets[t] >= n) fail(GLX_EINVAL, "target index out of range");
}

inline void fill_params(SegDesc& d, const glx_params& p, uint32_t n) {
  d.n = n;
  d.g = (float)p.g;
  d.eps2 = (float)(p.softening * p.softening);
  d.theta = (float)p.theta;
  d.precision = p.precision;
  d.leaf_size = p.leaf_size ? p.leaf_size : 1u;
  d.group_size = p.group_size ? p.group_size : (uint32_t)kGroup;
  d.group_lanes = p.group_lanes ? p.group_lanes : 4u;
  d.far2 = (float)(p.far_theta * p.far_theta);
  d.mono_inv = p.mono_ratio > 0.0 ? (float)(1.0 / (p.theta * p.mono_ratio)) : 0.0f;
}

inline hipStream_t make_stream() {
  hipStream_t st = nullptr;
  check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  return st;
}

// RAII device buffer for one-shot calls.
template <class T>
struct DevArray {
  T* p = nullptr;
  explicit DevArray(size_t n) {
    if (n) dev_malloc(reinterpret_cast<void**>(&p), n * sizeof(T), "hipMalloc");
  }
  ~DevArray() { dev_free(p); }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
};

struct Stream {
  hipStream_t st;
  Stream() : st(make_stream()) {}
  ~Stream() { (void)hipStreamDestroy(st); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
};

inline void h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "upload");
}
inline void d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "download");
}
inline void sync(hipStream_t st) { check(hipStreamSynchronize(st), "hipStreamSynchronize"); }

// Zero the walk statistics (not the bottom-up root counter, which the kernels keep
// at zero between builds).
inline void zero_walk_counters(const SegDesc& d, hipStream_t st) {
  check(hipMemsetAsync(d.counters + kCntInteractions, 0, 2 * sizeof(unsigned long long), st),
        "memset counters");
  check(hipMemsetAsync(d.counters + kCntOverflow, 0, 5 * sizeof(unsigned long long), st),
        "memset counters");
}

inline glx_stats stats_from(const unsigned long long* c) {
  glx_stats s{};
  s.point_terms = c[kCntInteractions];
  s.quad_terms = c[kCntQuad];
  s.node_tests = c[kCntVisits];
  s.fallbacks = c[kCntOverflow];
  s.max_stack = c[kCntMaxStack];
  s.far_nodes = c[kCntFar];
  s.rounds = c[kCntRounds];
  return s;
}

// Upload one system's state: f64 positions and masses as given, velocities narrowed
// to f32 on the host.
inline void upload_state(const SegDesc& d, size_t n, const double* pos, const double* vel,
                         const double* mass, hipStream_t st) {
  std::vector<float> v32(3 * n);
  for (size_t k = 0; k < 3 * n; ++k) v32[k] = (float)vel[k];
  h2d(d.x, pos, 3 * n * sizeof(double), st);
  h2d(d.v, v32.data(), 3 * n * sizeof(float), st);
  h2d(d.m, mass, n * sizeof(double), st);
  sync(st);  // v32 goes out of scope
}

// Phases of one resident KDK step, in launch order.
enum { kPhKick = 0, kPhEncode, kPhSort, kPhGather, kPhKarras, kPhAggregate, kPhWalk, kPhases };

// k KDK steps of the batch with direct launches and a HIP event between phases on the
// batch stream; accumulates device ms per phase into phase_ms[kPhases].
inline void timed_steps(Batch& b, uint64_t k, double* phase_ms) {
  hipStream_t st = b.stream();
  const size_t ne = (size_t)k * (kPhases + 1);
  std::vector<hipEvent_t> ev(ne, nullptr);
  struct EvGuard {
    std::vector<hipEvent_t>& e;
    ~EvGuard() {
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } eg{ev};
  for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
  for (uint64_t q = 0; q < k; ++q) {
    hipEvent_t* e = &ev[q * (kPhases + 1)];
    check(hipEventRecord(e[0], st), "hipEventRecord");
    b.kick_drift_bbox();
    check(hipEventRecord(e[1], st), "hipEventRecord");
    b.encode();
    check(hipEventRecord(e[2], st), "hipEventRecord");
    b.sort();
    check(hipEventRecord(e[3], st), "hipEventRecord");
    b.gather();
    check(hipEventRecord(e[4], st), "hipEventRecord");
    b.karras();
    check(hipEventRecord(e[5], st), "hipEventRecord");
    b.aggregate();
    check(hipEventRecord(e[6], st), "hipEventRecord");
    b.walk(1, false);
    check(hipEventRecord(e[7], st), "hipEventRecord");
  }
  sync(st);
  for (uint64_t q = 0; q < k; ++q) {
    hipEvent_t* e = &ev[q * (kPhases + 1)];
    for (int ph = 0; ph < kPhases; ++ph) {
      float ms = 0.f;
      check(hipEventElapsedTime(&ms, e[ph], e[ph + 1]), "hipEventElapsedTime");
      phase_ms[ph] += ms;
    }
  }
}

// One KDK step of the batch, captured as a hipGraph (kernels only).
struct StepGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~StepGraph() { drop(); }
  void capture(Batch& b) {
    std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());
    drop();
    hipStream_t st = b.stream();
    check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "BeginCapture");
    try {
      b.kick_drift_bbox();
      b.build_and_walk(1, false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(st, &graph), "EndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "GraphInstantiate");
  }
  void launch(hipStream_t st) { check(hipGraphLaunch(exec, st), "GraphLaunch"); }
};

}  // namespace abi
}  // namespace glx
