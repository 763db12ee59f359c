// glx_direct_sum_f64_noalloc: the FP64 direct sum of glx_direct_sum_f64 (same kernels,
// same chunk plan, so the same bits) with no device allocation. Everything it needs on
// the device beyond the caller's host arrays (staged positions, masses and targets, the
// output, the per-chunk partial sums) lives in one static device array of 64 MiB per
// GPU, loaded with the code object and held until the process ends. Its device address
// comes from hipGetSymbolAddress, once per device, on the first call on that device,
// which also creates the device's stream for these calls; the array's name is never
// passed to a kernel or a copy. Calls on one device are
// serialised by a mutex (they share the array); the array keeps the last call's
// contents and every byte a call reads is written by that call first.
#include <mutex>
#include <string>

#include "glx_abi_util.h"

using namespace glx;
using namespace glx::abi;

constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr size_t kNoAllocStageBytes = kNoAllocBytes - kDirectPartialBytes;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
struct NoAllocDevice {
  std::once_flag once;
  int status = 0;
  std::string msg;
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  hipStream_t st = nullptr;      // this device's stream for the calls (non-blocking)
  std::mutex mu;
};
NoAllocDevice g_noalloc[kNoAllocDevices];

inline size_t up256(size_t v) { return (v + 255) / 256 * 256; }
}  // namespace

// Staging bytes of one call: x (3n f64), m (n f64), targets (k u32), out (3k f64).
static size_t noalloc_stage_bytes(size_t n, size_t k) {
  return up256(3 * n * sizeof(double)) + up256(n * sizeof(double)) +
         up256(k * sizeof(uint32_t)) + up256(3 * k * sizeof(double));
}

extern "C" int glx_direct_sum_f64_noalloc(size_t n, const double* pos, const double* mass,
                                          double g, double softening, const uint32_t* targets,
                                          size_t n_targets, double* acc_out) {
  return guard([&] {
    validate_direct(n, pos, mass, targets, n_targets, acc_out);
    if (n_targets == 0) return;
    if (noalloc_stage_bytes(n, n_targets) > kNoAllocStageBytes)
      fail(GLX_EINVAL, "glx_direct_sum_f64_noalloc: 32 n + 28 n_targets bytes (rounded up to "
                       "256 per array) exceed its 48 MiB staging area");
    ensure_device();
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocDevices) fail(GLX_EDEVICE, "device index out of range");
    NoAllocDevice& D = g_noalloc[dev];
    std::call_once(D.once, [&D] {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        D.status = GLX_EDEVICE;
        D.msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
      }
      D.ptr = static_cast<unsigned char*>(p);
      // a stream of its own: work on the legacy null stream would depend on (and could
      // break) a graph capture in progress on another thread
      const hipError_t es = hipStreamCreateWithFlags(&D.st, hipStreamNonBlocking);
      if (es != hipSuccess) {
        D.status = GLX_EDEVICE;
        D.msg = std::string("hipStreamCreate: ") + hipGetErrorString(es);
      }
    });
    if (D.status) fail(D.status, D.msg);
    std::lock_guard<std::mutex> lock(D.mu);
    unsigned char* q = D.ptr;
    double* x = reinterpret_cast<double*>(q);
    q += up256(3 * n * sizeof(double));
    double* m = reinterpret_cast<double*>(q);
    q += up256(n * sizeof(double));
    uint32_t* tg = reinterpret_cast<uint32_t*>(q);
    q += up256(n_targets * sizeof(uint32_t));
    double* out = reinterpret_cast<double*>(q);
    double* part = reinterpret_cast<double*>(D.ptr + kNoAllocStageBytes);
    const DirectPlan plan = direct_plan((uint32_t)n, (uint32_t)n_targets, kDirectPartialBytes);
    hipStream_t st = D.st;
    h2d(x, pos, 3 * n * sizeof(double), st);
    h2d(m, mass, n * sizeof(double), st);
    h2d(tg, targets, n_targets * sizeof(uint32_t), st);
    check(launch_direct_f64(st, x, m, (uint32_t)n, tg, (uint32_t)n_targets, g,
                            softening * softening, plan, part, out),
          "direct_f64");
    d2h(acc_out, out, 3 * n_targets * sizeof(double), st);
    sync(st);
  });
}
