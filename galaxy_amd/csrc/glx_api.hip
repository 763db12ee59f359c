// C ABI of the LBVH gravity step (include/galaxy_lbvh.h). Every entry point converts
// runtime failures into a GLX_* status plus a thread-local message.
#include <stdio.h>

#include "galaxy_sim.h"
#include "glx_abi_util.h"
#include "glx_ic_device.h"
#include "glx_snapshot.h"

using namespace glx;
using namespace glx::abi;

thread_local std::string glx::abi::last_error;

extern "C" const char* glx_last_error(void) { return last_error.c_str(); }

extern "C" int glx_device_info(char* name, size_t name_len, int* compute_units,
                               uint64_t* hbm_bytes) {
  return guard([&] {
    ensure_device();
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
    if (name && name_len) {  // the marketing name, else the gfx target (e.g. gfx950:...)
      strncpy(name, prop.name[0] ? prop.name : prop.gcnArchName, name_len - 1);
      name[name_len - 1] = 0;
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
  });
}

extern "C" int glx_device_malloc(size_t bytes, void** ptr) {
  return guard([&] {
    if (!ptr) fail(GLX_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    ensure_device();
    if (bytes == 0) return;
    std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());
    if (hipMalloc(ptr, bytes) != hipSuccess) {
      *ptr = nullptr;
      fail(GLX_ENOMEM, "hipMalloc failed");
    }
  });
}

extern "C" int glx_device_free(void* ptr) {
  return guard([&] {
    if (!ptr) return;
    std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());
    check(hipFree(ptr), "hipFree");
  });
}

extern "C" int glx_device_to_host(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (bytes && (!dst || !src)) fail(GLX_EINVAL, "NULL pointer");
    if (!bytes) return;
    std::lock_guard<std::recursive_mutex> lock(device_memory_mutex());  // a null-stream copy
    check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
  });
}

// ----------------------------------------------------------------------------------
// ForceSolver
// ----------------------------------------------------------------------------------

struct glx_solver {
  glx_params p{};
  hipStream_t st = nullptr;
  Segment seg;
  Batch* batch = nullptr;
  uint32_t n_cur = 0xFFFFFFFFu;
  glx_stats stats{};
  double* pot_buf = nullptr;  // potential_energy's own x[3n], m[n] and block partials
  size_t pot_cap = 0;         // doubles
  uint8_t* mask = nullptr;  // walk mask of the active calls (original order)
  size_t mask_cap = 0;
  uint32_t cache_n = 0;     // bodies of the tree kept by the last build (0: none)
  ~glx_solver() {
    delete batch;
    dev_free(pot_buf);
    dev_free(mask);
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" int glx_solver_create(const glx_params* params, glx_solver** out) {
  return guard([&] {
    if (!out) fail(GLX_EINVAL, "out is NULL");
    *out = nullptr;
    validate_params(params);
    ensure_device();
    glx_solver* s = new glx_solver();
    try {
      s->p = *params;
      s->st = make_stream();
      s->batch = new Batch(s->st);
      s->batch->set_mode(mode_from(s->p));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

extern "C" int glx_solver_destroy(glx_solver* s) {
  return guard([&] { delete s; });
}

// Upload positions/masses into the solver's segment and (re)describe it when its
// size or storage changed.
static void solver_stage_inputs(glx_solver* s, size_t n, const double* pos, const double* mass) {
  const uint32_t cap_before = s->seg.capacity;
  s->seg.reserve((uint32_t)n, true);
  fill_params(s->seg.desc, s->p, (uint32_t)n);
  if (s->n_cur != n || s->seg.capacity != cap_before) {
    s->batch->set_segments({s->seg.desc});
    s->n_cur = (uint32_t)n;
    s->cache_n = 0;  // the kept tree described other storage or another n
  }
  h2d(s->seg.desc.x, pos, 3 * n * sizeof(double), s->st);
  h2d(s->seg.desc.m, mass, n * sizeof(double), s->st);
}

static void solver_eval(glx_solver* s, size_t n, const double* pos, const double* mass,
                        const size_t* active, size_t n_active, double* acc, bool cached);

extern "C" int glx_solver_accelerations(glx_solver* s, size_t n, const double* pos,
                                        const double* mass, double* acc) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "solver is NULL");
    validate_n(n);
    if (n == 0) return;
    if (!pos || !mass || !acc) fail(GLX_EINVAL, "NULL array");
    s->stats = glx_stats{};
    if (n == 1) {
      acc[0] = acc[1] = acc[2] = 0.0;
      return;
    }
    solver_eval(s, n, pos, mass, nullptr, 0, acc, false);
  });
}

// Uploads the walk mask of `active` (original indices), or clears it for a full walk.
static void solver_set_mask(glx_solver* s, size_t n, const size_t* active, size_t n_active) {
  if (!active) {
    s->batch->set_active(nullptr);
    return;
  }
  std::vector<uint8_t> m(n, 0);
  for (size_t q = 0; q < n_active; ++q) {
    if (active[q] >= n) fail(GLX_EINVAL, "active index out of range");
    m[active[q]] = 1;
  }
  if (s->mask_cap < n) {
    dev_free(s->mask);
    s->mask = nullptr;
    s->mask_cap = 0;
    dev_malloc(reinterpret_cast<void**>(&s->mask), n, "hipMalloc");
    s->mask_cap = n;
  }
  h2d(s->mask, m.data(), n, s->st);
  s->batch->set_active(s->mask);
}

// One force evaluation of n >= 2 bodies: a fresh build (kept as the cached tree) or, with
// `cached`, the kept tree with the leaves moved to the new positions; walks every group
// (active == null) or the groups holding an active body.
static void solver_eval(glx_solver* s, size_t n, const double* pos, const double* mass,
                        const size_t* active, size_t n_active, double* acc, bool cached) {
  if (cached && s->cache_n != n) fail(GLX_EINVAL, "no cached tree of n bodies");
  if (cached) {
    h2d(s->seg.desc.x, pos, 3 * n * sizeof(double), s->st);
    h2d(s->seg.desc.m, mass, n * sizeof(double), s->st);
  } else {
    s->cache_n = 0;
    solver_stage_inputs(s, n, pos, mass);
  }
  const bool count = (s->p.flags & GLX_FLAG_COUNT) != 0;
  if (count) zero_walk_counters(s->seg.desc, s->st);
  Batch& b = *s->batch;
  solver_set_mask(s, n, active, n_active);
  if (cached) {
    b.refresh_leaves();
    b.walk(0, count);
  } else {
    b.reset_bbox();
    b.bbox();
    b.build_and_walk(0, count);
  }
  d2h(acc, s->seg.desc.acc_out, 3 * n * sizeof(double), s->st);
  unsigned long long c[kCounters] = {};
  if (count) d2h(c, s->seg.desc.counters, sizeof(c), s->st);
  sync(s->st);
  if (active) b.set_active(nullptr);
  s->stats = stats_from(c);
  s->cache_n = (uint32_t)n;
}

extern "C" int glx_solver_accelerations_active(glx_solver* s, size_t n, const double* pos,
                                               const double* mass, const size_t* active,
                                               size_t n_active, double* acc) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "solver is NULL");
    validate_n(n);
    if (n == 0) return;
    if (!pos || !mass || !acc || (n_active && !active)) fail(GLX_EINVAL, "NULL array");
    s->stats = glx_stats{};
    if (n == 1) {
      acc[0] = acc[1] = acc[2] = 0.0;
      return;
    }
    solver_eval(s, n, pos, mass, active ? active : nullptr, n_active, acc, false);
  });
}

extern "C" int glx_solver_rebuild_cache(glx_solver* s, size_t n, const double* pos,
                                        const double* mass) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "solver is NULL");
    validate_n(n);
    s->cache_n = 0;
    if (n < 2) {
      s->cache_n = (uint32_t)n;
      return;
    }
    if (!pos || !mass) fail(GLX_EINVAL, "NULL array");
    solver_stage_inputs(s, n, pos, mass);
    Batch& b = *s->batch;
    b.reset_bbox();
    b.bbox();
    b.encode();
    b.sort();
    b.gather();
    b.karras();
    b.aggregate();
    sync(s->st);
    s->cache_n = (uint32_t)n;
  });
}

extern "C" int glx_solver_gravity_active_cached(glx_solver* s, size_t n, const double* pos,
                                                const double* mass, const size_t* active,
                                                size_t n_active, double* acc) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "solver is NULL");
    validate_n(n);
    if (n != s->cache_n) fail(GLX_EINVAL, "no cached tree of n bodies");
    if (n == 0) return;
    if (!pos || !mass || !acc || (n_active && !active)) fail(GLX_EINVAL, "NULL array");
    s->stats = glx_stats{};
    if (n == 1) {
      acc[0] = acc[1] = acc[2] = 0.0;
      return;
    }
    solver_eval(s, n, pos, mass, active, n_active, acc, true);
  });
}

extern "C" int glx_solver_stats(glx_solver* s, glx_stats* out) {
  return guard([&] {
    if (!s || !out) fail(GLX_EINVAL, "NULL argument");
    *out = s->stats;
  });
}

extern "C" int glx_solver_potential_energy(glx_solver* s, size_t n, const double* pos,
                                           const double* mass, double* out) {
  return guard([&] {
    if (!s || !out) fail(GLX_EINVAL, "NULL argument");
    validate_n(n);
    *out = 0.0;
    if (n < 2) return;
    if (!pos || !mass) fail(GLX_EINVAL, "NULL array");
    // The potential works on its own copy of the bodies, so the solver's segment and
    // the tree kept for gravity_active_cached stay as the last force evaluation left
    // them (the reference's potential_energy takes &self).
    const size_t nb = (n + 255) / 256 + 1;
    const size_t need = 4 * n + nb;
    if (need > s->pot_cap) {
      dev_free(s->pot_buf);
      s->pot_buf = nullptr;
      s->pot_cap = 0;
      dev_malloc(reinterpret_cast<void**>(&s->pot_buf), need * sizeof(double),
                 "hipMalloc(potential)");
      s->pot_cap = need;
    }
    double* px = s->pot_buf;
    double* pm = px + 3 * n;
    double* part = pm + n;
    h2d(px, pos, 3 * n * sizeof(double), s->st);
    h2d(pm, mass, n * sizeof(double), s->st);
    check(launch_potential_f64(s->st, px, pm, (uint32_t)n, s->p.g,
                               s->p.softening * s->p.softening, part, part + (nb - 1)),
          "potential");
    d2h(out, part + (nb - 1), sizeof(double), s->st);
    sync(s->st);
  });
}

extern "C" int glx_direct_sum_f64(size_t n, const double* pos, const double* mass, double g,
                                  double softening, const uint32_t* targets, size_t n_targets,
                                  double* acc_out) {
  return guard([&] {
    validate_direct(n, pos, mass, targets, n_targets, acc_out);
    if (n_targets == 0) return;
    ensure_device();
    const DirectPlan plan = direct_plan((uint32_t)n, (uint32_t)n_targets, kDirectPartialBytes);
    Stream S;
    DevArray<double> x(3 * n), m(n), out(3 * n_targets);
    DevArray<double> part(plan.chunks > 1 ? 3 * n_targets * plan.chunks : 0);
    DevArray<uint32_t> tg(n_targets);
    h2d(x.p, pos, 3 * n * sizeof(double), S.st);
    h2d(m.p, mass, n * sizeof(double), S.st);
    h2d(tg.p, targets, n_targets * sizeof(uint32_t), S.st);
    check(launch_direct_f64(S.st, x.p, m.p, (uint32_t)n, tg.p, (uint32_t)n_targets, g,
                            softening * softening, plan, part.p, out.p),
          "direct_f64");
    d2h(acc_out, out.p, 3 * n_targets * sizeof(double), S.st);
    sync(S.st);
  });
}

// ----------------------------------------------------------------------------------
// stage exports
// ----------------------------------------------------------------------------------

namespace {
struct StageCtx {
  Stream S;
  Segment seg;
  Batch batch;
  explicit StageCtx(bool key63 = false, bool quad = false, int radix_bits = kRadixBits)
      : batch(S.st) {
    Mode m;
    m.key63 = key63;
    m.quad = quad;
    m.radix_bits = radix_bits;
    batch.set_mode(m);
  }
  void describe(uint32_t n) {
    seg.reserve(n, false);
    glx_params p{};
    p.g = 1.0;
    p.softening = 0.0;
    p.theta = 1.0;
    fill_params(seg.desc, p, n);
    batch.set_segments({seg.desc});
  }
};

template <class K>
void stage_morton(size_t n, const double* pos, double* bounds_out, uint32_t* lanes, K* codes) {
  validate_n(n);
  if (n == 0) fail(GLX_EINVAL, "reference_morton requires a non-empty system");
  if (!pos || !codes) fail(GLX_EINVAL, "NULL array");
  ensure_device();
  StageCtx c(sizeof(K) == 8);
  c.describe((uint32_t)n);
  DevArray<uint32_t> dl(lanes ? 3 * n : 0);
  h2d(c.seg.desc.x, pos, 3 * n * sizeof(double), c.S.st);
  c.batch.reset_bbox();
  c.batch.bbox();
  c.batch.encode(dl.p);
  d2h(codes, c.seg.desc.keys[0], n * sizeof(K), c.S.st);
  if (lanes) d2h(lanes, dl.p, 3 * n * sizeof(uint32_t), c.S.st);
  double b[8];
  d2h(b, c.seg.desc.bounds, sizeof(b), c.S.st);
  sync(c.S.st);
  if (bounds_out) memcpy(bounds_out, b, 5 * sizeof(double));
}

template <class K>
void stage_sort(size_t n, const K* codes, uint32_t* order, K* sorted_codes,
                uint32_t radix_bits = kRadixBits) {
  validate_n(n);
  if (radix_bits < 8 || radix_bits > (uint32_t)kRadixBitsMax)
    fail(GLX_EINVAL, "radix_bits must be 8, 9 or 10");
  if (n == 0) return;
  if (!codes) fail(GLX_EINVAL, "NULL array");
  constexpr bool key63 = sizeof(K) == 8;
  const int span = sort_passes_for(key63, (int)radix_bits) * (int)radix_bits;
  if (span < (int)(8 * sizeof(K)))
    for (size_t k = 0; k < n; ++k)
      if (codes[k] >> span) fail(GLX_EINVAL, "a code exceeds the sort passes' bits");
  ensure_device();
  StageCtx c(key63, false, (int)radix_bits);
  c.describe((uint32_t)n);
  h2d(c.seg.desc.keys[sort_start(key63, (int)radix_bits)], codes, n * sizeof(K), c.S.st);
  c.batch.iota();
  c.batch.sort();
  if (order) d2h(order, c.seg.desc.vals[0], n * sizeof(uint32_t), c.S.st);
  if (sorted_codes) d2h(sorted_codes, c.seg.desc.keys[0], n * sizeof(K), c.S.st);
  sync(c.S.st);
}

template <class K>
void stage_karras(size_t n, const K* sorted_codes, uint32_t* left, uint32_t* right,
                  uint32_t* parent) {
  validate_n(n);
  if (n == 0) fail(GLX_EINVAL, "reference_karras requires a non-empty system");
  if (!sorted_codes || !parent || (n > 1 && (!left || !right))) fail(GLX_EINVAL, "NULL array");
  for (size_t k = 1; k < n; ++k)
    if (sorted_codes[k - 1] > sorted_codes[k]) fail(GLX_EINVAL, "codes must be sorted");
  if (n == 1) {
    parent[0] = kNoParent;
    return;
  }
  ensure_device();
  StageCtx c(sizeof(K) == 8);
  c.describe((uint32_t)n);
  h2d(c.seg.desc.keys[0], sorted_codes, n * sizeof(K), c.S.st);
  c.batch.karras();
  std::vector<uint4> topo(n - 1);
  d2h(topo.data(), c.seg.desc.topo, (n - 1) * sizeof(uint4), c.S.st);
  d2h(parent, c.seg.desc.parent_leaf, n * sizeof(uint32_t), c.S.st);
  sync(c.S.st);
  const uint32_t nn = (uint32_t)n;
  for (size_t j = 0; j + 1 < n; ++j) {
    left[j] = left_child(topo[j].z, nn);
    right[j] = right_child(topo[j].z, nn);
    parent[n + j] = topo[j].w;
  }
}
}  // namespace

extern "C" int glx_stage_morton(size_t n, const double* pos, double* bounds_out, uint32_t* lanes,
                                uint32_t* codes) {
  return guard([&] { stage_morton<uint32_t>(n, pos, bounds_out, lanes, codes); });
}

extern "C" int glx_stage_morton63(size_t n, const double* pos, double* bounds_out,
                                  uint32_t* lanes, uint64_t* codes) {
  return guard([&] { stage_morton<uint64_t>(n, pos, bounds_out, lanes, codes); });
}

extern "C" int glx_stage_sort(size_t n, const uint32_t* codes, uint32_t* order,
                              uint32_t* sorted_codes) {
  return guard([&] { stage_sort<uint32_t>(n, codes, order, sorted_codes); });
}

extern "C" int glx_stage_sort63(size_t n, const uint64_t* codes, uint32_t* order,
                                uint64_t* sorted_codes) {
  return guard([&] { stage_sort<uint64_t>(n, codes, order, sorted_codes); });
}

extern "C" int glx_stage_sort_bits(size_t n, const uint32_t* codes, uint32_t* order,
                                   uint32_t* sorted_codes, uint32_t radix_bits) {
  return guard([&] { stage_sort<uint32_t>(n, codes, order, sorted_codes, radix_bits); });
}

extern "C" int glx_stage_sort63_bits(size_t n, const uint64_t* codes, uint32_t* order,
                                     uint64_t* sorted_codes, uint32_t radix_bits) {
  return guard([&] { stage_sort<uint64_t>(n, codes, order, sorted_codes, radix_bits); });
}

extern "C" int glx_stage_karras(size_t n, const uint32_t* sorted_codes, uint32_t* left,
                                uint32_t* right, uint32_t* parent) {
  return guard([&] { stage_karras<uint32_t>(n, sorted_codes, left, right, parent); });
}

extern "C" int glx_stage_karras63(size_t n, const uint64_t* sorted_codes, uint32_t* left,
                                  uint32_t* right, uint32_t* parent) {
  return guard([&] { stage_karras<uint64_t>(n, sorted_codes, left, right, parent); });
}

extern "C" int glx_stage_aggregate(size_t n, const uint32_t* sorted_codes, const double* sorted_pos,
                                   const double* sorted_mass, double* origin_out, float* aabb_min,
                                   float* aabb_max, float* com, float* mass, float* quad) {
  return guard([&] {
    validate_n(n);
    if (n == 0) fail(GLX_EINVAL, "reference_aggregate requires a non-empty system");
    if (!sorted_codes || !sorted_pos || !sorted_mass || !aabb_min || !aabb_max || !com || !mass)
      fail(GLX_EINVAL, "NULL array");
    for (size_t k = 1; k < n; ++k)
      if (sorted_codes[k - 1] > sorted_codes[k]) fail(GLX_EINVAL, "codes must be sorted");
    ensure_device();
    StageCtx c(false, quad != nullptr);
    c.describe((uint32_t)n);
    h2d(c.seg.desc.x, sorted_pos, 3 * n * sizeof(double), c.S.st);
    h2d(c.seg.desc.m, sorted_mass, n * sizeof(double), c.S.st);
    h2d(c.seg.desc.keys[0], sorted_codes, n * sizeof(uint32_t), c.S.st);
    c.batch.reset_bbox();
    c.batch.bbox();
    c.batch.bounds_only();
    c.batch.iota();
    c.batch.gather();
    c.batch.karras();
    c.batch.aggregate();
    std::vector<float4> lf(n), qd(quad && n > 1 ? 2 * (n - 1) : 0);
    std::vector<Node> nd(n > 1 ? n - 1 : 0);
    std::vector<NodeBox> bx(n > 1 ? n - 1 : 0);
    double b[8];
    d2h(lf.data(), c.seg.desc.leaf, n * sizeof(float4), c.S.st);
    if (n > 1) {
      d2h(nd.data(), c.seg.desc.node, (n - 1) * sizeof(Node), c.S.st);
      d2h(bx.data(), c.seg.desc.box, (n - 1) * sizeof(NodeBox), c.S.st);
    }
    if (!qd.empty()) d2h(qd.data(), c.seg.desc.quad, qd.size() * sizeof(float4), c.S.st);
    d2h(b, c.seg.desc.bounds, sizeof(b), c.S.st);
    sync(c.S.st);
    if (origin_out) memcpy(origin_out, b + 5, 3 * sizeof(double));
    for (size_t k = 0; k < n; ++k) {
      const float p[3] = {lf[k].x, lf[k].y, lf[k].z};
      for (int a = 0; a < 3; ++a) aabb_min[3 * k + a] = aabb_max[3 * k + a] = com[3 * k + a] = p[a];
      mass[k] = lf[k].w;
      if (quad)
        for (int w = 0; w < 6; ++w) quad[6 * k + w] = 0.f;
    }
    for (size_t j = 0; j + 1 < n; ++j) {
      const size_t u = n + j;
      const Node& q = nd[j];
      const NodeBox& w = bx[j];
      const float cm[3] = {q.a.x, q.a.y, q.a.z}, mn[3] = {w.lo.x, w.lo.y, w.lo.z},
                  mx[3] = {w.hi.x, w.hi.y, w.hi.z};
      for (int a = 0; a < 3; ++a) {
        com[3 * u + a] = cm[a];
        aabb_min[3 * u + a] = mn[a];
        aabb_max[3 * u + a] = mx[a];
      }
      mass[u] = q.a.w;
      if (quad) {
        const float qq[6] = {qd[2 * j].x, qd[2 * j].y, qd[2 * j].z,
                             qd[2 * j].w, qd[2 * j + 1].x, qd[2 * j + 1].y};
        for (int w = 0; w < 6; ++w) quad[6 * u + w] = qq[w];
      }
    }
  });
}

extern "C" int glx_stage_flatten(size_t n, const uint32_t* sorted_codes, const double* sorted_pos,
                                 const double* sorted_mass, double* origin_out, uint32_t* next,
                                 uint32_t* body_start, uint32_t* body_count, float* center,
                                 float* half, float* com, float* mass, float* delta) {
  return guard([&] {
    validate_n(n);
    if (n == 0) fail(GLX_EINVAL, "reference_flatten requires a non-empty system");
    if (!sorted_codes || !sorted_pos || !sorted_mass || !next || !body_start || !body_count ||
        !center || !half || !com || !mass || !delta)
      fail(GLX_EINVAL, "NULL array");
    for (size_t k = 1; k < n; ++k)
      if (sorted_codes[k - 1] > sorted_codes[k]) fail(GLX_EINVAL, "codes must be sorted");
    ensure_device();
    StageCtx c(false, false);
    c.describe((uint32_t)n);
    h2d(c.seg.desc.x, sorted_pos, 3 * n * sizeof(double), c.S.st);
    h2d(c.seg.desc.m, sorted_mass, n * sizeof(double), c.S.st);
    h2d(c.seg.desc.keys[0], sorted_codes, n * sizeof(uint32_t), c.S.st);
    c.batch.reset_bbox();
    c.batch.bbox();
    c.batch.bounds_only();
    c.batch.iota();
    c.batch.gather();
    c.batch.karras();
    c.batch.aggregate();
    const size_t t = 2 * n - 1;
    DevArray<uint32_t> u32(3 * t);
    DevArray<float> f32(11 * t);
    const FlatOut o{u32.p, u32.p + t, u32.p + 2 * t, f32.p, f32.p + 3 * t,
                    f32.p + 6 * t, f32.p + 9 * t, f32.p + 10 * t};
    c.batch.flatten(o);
    d2h(next, o.next, t * sizeof(uint32_t), c.S.st);
    d2h(body_start, o.body_start, t * sizeof(uint32_t), c.S.st);
    d2h(body_count, o.body_count, t * sizeof(uint32_t), c.S.st);
    d2h(center, o.center, 3 * t * sizeof(float), c.S.st);
    d2h(half, o.half, 3 * t * sizeof(float), c.S.st);
    d2h(com, o.com, 3 * t * sizeof(float), c.S.st);
    d2h(mass, o.mass, t * sizeof(float), c.S.st);
    d2h(delta, o.delta, t * sizeof(float), c.S.st);
    double b[8];
    d2h(b, c.seg.desc.bounds, sizeof(b), c.S.st);
    sync(c.S.st);
    if (origin_out) memcpy(origin_out, b + 5, 3 * sizeof(double));
  });
}

// ----------------------------------------------------------------------------------
// resident stepper
// ----------------------------------------------------------------------------------

struct glx_sys {
  glx_params p{};
  hipStream_t st = nullptr;
  Segment seg;
  Batch* batch = nullptr;
  uint32_t n = 0;
  double time = 0.0;
  double cur_dt = NAN;
  StepGraph graph;
  ~glx_sys() {
    graph.drop();
    delete batch;
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" int glx_sys_create(const glx_params* params, glx_sys** out) {
  return guard([&] {
    if (!out) fail(GLX_EINVAL, "out is NULL");
    *out = nullptr;
    validate_params(params);
    ensure_device();
    glx_sys* s = new glx_sys();
    try {
      s->p = *params;
      s->st = make_stream();
      s->batch = new Batch(s->st);
      s->batch->set_mode(mode_from(s->p));
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

extern "C" int glx_sys_destroy(glx_sys* s) {
  return guard([&] { delete s; });
}

// Replace the system's state with n particles written by load(desc) on the system's
// stream, prime a(x0) and reset the time.
template <class Load>
static void sys_load(glx_sys* s, size_t n, Load&& load) {
  s->graph.drop();
  s->time = 0.0;
  s->n = 0;
  s->cur_dt = NAN;
  if (n == 0) return;
  s->seg.reserve((uint32_t)n, false);
  fill_params(s->seg.desc, s->p, (uint32_t)n);
  s->seg.desc.dt = 0.0;
  s->seg.desc.half_dt = 0.0f;
  s->batch->set_segments({s->seg.desc});
  load(s->seg.desc);
  Batch& b = *s->batch;
  b.reset_bbox();
  b.bbox();
  b.build_and_walk(0, false);  // prime a(x0)
  sync(s->st);
  s->n = (uint32_t)n;
}

extern "C" int glx_sys_upload(glx_sys* s, size_t n, const double* pos, const double* vel,
                              const double* mass) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "sys is NULL");
    validate_n(n);
    if (n && (!pos || !vel || !mass)) fail(GLX_EINVAL, "NULL array");
    sys_load(s, n, [&](const SegDesc& d) { upload_state(d, n, pos, vel, mass, s->st); });
  });
}

extern "C" int glx_sys_generate(glx_sys* s, const glx_ic_spec* spec) {
  return guard([&] {
    if (!s || !spec) fail(GLX_EINVAL, "NULL argument");
    ic::validate_spec(*spec);
    sys_load(s, spec->n,
             [&](const SegDesc& d) { ic::generate_resident(*spec, d.x, d.v, d.m, s->st); });
  });
}

// k KDK steps of dt through the captured step graph (re-captured when dt changes).
static void sys_step(glx_sys* s, double dt, uint64_t k) {
  if (!isfinite(dt)) fail(GLX_EINVAL, "dt must be finite");
  if (s->n == 0) {
    s->time += (double)k * dt;
    return;
  }
  if (k == 0) return;
  if (!(dt == s->cur_dt)) {
    s->batch->set_dt(dt);
    s->cur_dt = dt;
  }
  if (!s->graph.exec) s->graph.capture(*s->batch);
  for (uint64_t q = 0; q < k; ++q) {
    s->graph.launch(s->st);
    s->time += dt;
  }
  sync(s->st);
}

extern "C" int glx_sys_step(glx_sys* s, double dt, uint64_t k) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "sys is NULL");
    sys_step(s, dt, k);
  });
}

extern "C" int glx_sys_step_timed(glx_sys* s, double dt, uint64_t k, double* phase_ms) {
  return guard([&] {
    if (!s || !phase_ms) fail(GLX_EINVAL, "NULL argument");
    if (!isfinite(dt)) fail(GLX_EINVAL, "dt must be finite");
    for (int q = 0; q < kPhases; ++q) phase_ms[q] = 0.0;
    if (s->n == 0 || k == 0) {
      s->time += (double)k * dt;
      return;
    }
    if (!(dt == s->cur_dt)) {
      s->batch->set_dt(dt);
      s->cur_dt = dt;
    }
    timed_steps(*s->batch, k, phase_ms);
    s->time += (double)k * dt;
  });
}

extern "C" int glx_sys_count(glx_sys* s, glx_stats* out) {
  return guard([&] {
    if (!s || !out) fail(GLX_EINVAL, "NULL argument");
    unsigned long long c[kCounters] = {};
    if (s->n > 1) {
      Batch& b = *s->batch;
      zero_walk_counters(s->seg.desc, s->st);
      b.reset_bbox();
      b.bbox();
      b.build_and_walk(0, true);
      d2h(c, s->seg.desc.counters, sizeof(c), s->st);
      sync(s->st);
    }
    *out = stats_from(c);
  });
}

extern "C" int glx_sys_snapshot(glx_sys* s, double* pos, double* vel) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "sys is NULL");
    const size_t n = s->n;
    if (n == 0) return;
    std::vector<float> v32(vel ? 3 * n : 0);
    if (pos) d2h(pos, s->seg.desc.x, 3 * n * sizeof(double), s->st);
    if (vel) d2h(v32.data(), s->seg.desc.v, 3 * n * sizeof(float), s->st);
    sync(s->st);
    if (vel)
      for (size_t q = 0; q < 3 * n; ++q) vel[q] = (double)v32[q];
  });
}

extern "C" int glx_sys_snapshot_accel(glx_sys* s, double* acc) {
  return guard([&] {
    if (!s || !acc) fail(GLX_EINVAL, "NULL argument");
    const size_t n = s->n;
    if (n == 0) return;
    std::vector<float> a32(3 * n);
    d2h(a32.data(), s->seg.desc.acc, 3 * n * sizeof(float), s->st);
    sync(s->st);
    for (size_t q = 0; q < 3 * n; ++q) acc[q] = (double)a32[q];
  });
}

extern "C" int glx_sys_snapshot_mass(glx_sys* s, double* mass) {
  return guard([&] {
    if (!s || !mass) fail(GLX_EINVAL, "NULL argument");
    if (s->n == 0) return;
    d2h(mass, s->seg.desc.m, (size_t)s->n * sizeof(double), s->st);
    sync(s->st);
  });
}

extern "C" int glx_sys_time(glx_sys* s, double* t) {
  return guard([&] {
    if (!s || !t) fail(GLX_EINVAL, "NULL argument");
    *t = s->time;
  });
}

extern "C" int glx_sys_len(glx_sys* s, size_t* n) {
  return guard([&] {
    if (!s || !n) fail(GLX_EINVAL, "NULL argument");
    *n = s->n;
  });
}

extern "C" int glx_sys_diag(glx_sys* s, int with_potential, glx_diag* out) {
  return guard([&] {
    if (!s || !out) fail(GLX_EINVAL, "NULL argument");
    memset(out, 0, sizeof(*out));
    if (s->n == 0) return;
    double r[kDiagRecord];
    s->batch->diag(r);
    out->kinetic = r[0];
    for (int a = 0; a < 3; ++a) {
      out->momentum[a] = r[1 + a];
      out->angular_momentum[a] = r[4 + a];
      out->center_of_mass[a] = r[7 + a] / r[10];
    }
    out->mass = r[10];
    if (with_potential && s->n > 1) {
      const size_t nb = ((size_t)s->n + 255) / 256 + 1;
      DevArray<double> part(nb);
      check(launch_potential_f64(s->st, s->seg.desc.x, s->seg.desc.m, s->n, s->p.g,
                                 s->p.softening * s->p.softening, part.p, part.p + (nb - 1)),
            "potential");
      d2h(&out->potential, part.p + (nb - 1), sizeof(double), s->st);
      sync(s->st);
    }
  });
}

extern "C" int glx_sys_set_time(glx_sys* s, double time) {
  return guard([&] {
    if (!s) fail(GLX_EINVAL, "sys is NULL");
    if (!isfinite(time)) fail(GLX_EINVAL, "time must be finite");
    s->time = time;
  });
}

static const char kCodeVersion[] = "galaxy_amd 0.1.0";

// sim::run (sim/src/lib.rs:110-174): the state as uploaded is snapshot `first_step`, then
// steps advance in runs that end on the cadence (one sync per run), a snapshot follows
// every multiple of snapshot_every, and the final step is always written.
extern "C" int glx_sys_run(glx_sys* s, const glx_run_config* cfg, const char* dir,
                           const glx_snap_columns* cols, glx_run_summary* out) {
  return guard([&] {
    if (!s || !cfg || !dir) fail(GLX_EINVAL, "NULL argument");
    if (cfg->snapshot_every == 0) fail(GLX_EINVAL, "snapshot_every must be >= 1");
    if (!isfinite(cfg->dt) || !(cfg->dt > 0.0))
      fail(GLX_EINVAL, "dt must be a positive finite number");
    const size_t n = s->n;
    std::vector<double> pos(3 * n), vel(3 * n), mass(n);
    std::vector<float> v32(3 * n);
    if (n) {
      d2h(mass.data(), s->seg.desc.m, n * sizeof(double), s->st);
      sync(s->st);
    }
    uint64_t emitted = 0;
    auto emit = [&](uint64_t step) {
      if (n) {
        d2h(pos.data(), s->seg.desc.x, 3 * n * sizeof(double), s->st);
        d2h(v32.data(), s->seg.desc.v, 3 * n * sizeof(float), s->st);
        sync(s->st);
        for (size_t q = 0; q < 3 * n; ++q) vel[q] = (double)v32[q];
      }
      glx_snap_header h{};
      h.time = s->time;
      h.step = cfg->first_step + step;
      h.scale_factor = 1.0;
      h.softening = cfg->softening;
      h.rng_seed = cfg->rng_seed;
      h.config_hash = cfg->config_hash;
      h.units = cfg->units ? cfg->units : "";
      h.code_version = kCodeVersion;
      char name[48];
      snprintf(name, sizeof(name), "snapshot_%08llu.snap", (unsigned long long)h.step);
      const std::string path = std::string(dir) + "/" + name;
      glx::snap::write(path.c_str(), h, n, pos.data(), vel.data(), mass.data(), cols);
      ++emitted;
    };
    emit(0);
    uint64_t step = 0, last = 0;
    const uint64_t every = cfg->snapshot_every;
    while (step < cfg->n_steps) {
      const uint64_t cadence = (step / every + 1) * every;
      const uint64_t next = cadence < cfg->n_steps ? cadence : cfg->n_steps;
      sys_step(s, cfg->dt, next - step);
      step = next;
      if (step % every == 0) {
        emit(step);
        last = step;
      }
    }
    if (last != cfg->n_steps) emit(cfg->n_steps);
    if (out) {
      out->steps = cfg->n_steps;
      out->final_time = s->time;
      out->snapshots_emitted = emitted;
    }
  });
}
