// Batched and served forms: a pool keeps many systems ("jobs") resident on the device
// and advances all active ones by one KDK step per call with one hipGraph launch of the
// segment-aware kernels. Admission is bounded by a device-memory budget and a slot
// count; jobs join and leave between steps (continuous batching), and the step graph
// is re-captured whenever the set of active jobs changes.
#include <chrono>
#include <map>
#include <memory>

#include "glx_abi_util.h"
#include "glx_ic_device.h"

using namespace glx;
using namespace glx::abi;

namespace {
struct Job {
  uint64_t id = 0;
  Segment seg;
  uint64_t steps_target = 0;
  uint64_t steps_done = 0;
  double time = 0.0;
  bool done = false;
  bool evicted = false;  // paused by the step budget (until resumed or retired)
  uint64_t activated = 0;  // order of admission or last resumption
  bool primed = false;  // a(x0) computed (new jobs are primed together before their first step)
};
}  // namespace

struct glx_pool {
  glx_params build{};
  uint64_t budget = 0;
  uint32_t max_jobs = 0;
  hipStream_t st = nullptr;
  Batch* batch = nullptr;        // active jobs, stepped together
  Batch* prime = nullptr;        // one job at admission
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  std::vector<Job*> active;      // segment order of `batch`
  bool dirty = true;
  StepGraph graph;
  uint64_t next_id = 1;
  uint64_t next_activation = 1;
  uint64_t used = 0;
  uint64_t steps = 0;
  uint64_t evictions = 0;
  double step_budget_ms = 0.0;  // 0: no budget
  double refresh_ms = 0.0;      // host time of the last step's refresh
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~glx_pool() {
    graph.drop();
    jobs.clear();
    delete batch;
    delete prime;
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (st) (void)hipStreamDestroy(st);
  }
};

// Rebuild the stepping batch from the active jobs (id order), prime the jobs admitted
// since the last step with one batched force evaluation, and re-capture the step graph.
static void pool_refresh(glx_pool* p) {
  p->refresh_ms = 0.0;
  if (!p->dirty) return;
  const auto t0 = std::chrono::steady_clock::now();
  p->graph.drop();
  p->active.clear();
  std::vector<SegDesc> descs, fresh;
  std::vector<Job*> unprimed;
  for (auto& kv : p->jobs) {
    Job* j = kv.second.get();
    if (!j->primed) {
      fresh.push_back(j->seg.desc);
      unprimed.push_back(j);
    }
    if (j->done || j->evicted) continue;
    p->active.push_back(j);
    descs.push_back(j->seg.desc);
  }
  if (!fresh.empty()) {
    Batch& b = *p->prime;
    b.set_segments(fresh);
    b.reset_bbox();
    b.bbox();
    b.build_and_walk(0, false);
    for (Job* j : unprimed) j->primed = true;
  }
  if (!descs.empty()) {
    p->batch->set_segments(descs);
    p->graph.capture(*p->batch);
  }
  p->dirty = false;
  p->refresh_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// After a step over the budget: evict (pause) the most recently admitted or resumed
// active job.
static void pool_enforce_budget(glx_pool* p, double ms) {
  if (!(p->step_budget_ms > 0.0) || !(ms > p->step_budget_ms)) return;
  Job* last = nullptr;  // jobs that finished with this step are not evicted
  for (Job* j : p->active)
    if (!j->done && (!last || j->activated > last->activated)) last = j;
  if (!last) return;
  last->evicted = true;
  p->evictions += 1;
  p->dirty = true;
}

static Job* find_job(glx_pool* p, uint64_t id) {
  auto it = p->jobs.find(id);
  if (it == p->jobs.end()) fail(GLX_ENOTFOUND, "unknown job id");
  return it->second.get();
}

// After a step: advance bookkeeping, retire finished jobs from the active set.
static void pool_advance(glx_pool* p, uint64_t k) {
  for (Job* j : p->active) {
    j->steps_done += k;
    j->time += (double)k * j->seg.desc.dt;
    if (j->steps_target && j->steps_done >= j->steps_target) {
      j->done = true;
      p->dirty = true;
    }
  }
  p->steps += k;
}

extern "C" int glx_pool_create(const glx_params* build, uint64_t budget_bytes, uint32_t max_jobs,
                               glx_pool** out) {
  return guard([&] {
    if (!out) fail(GLX_EINVAL, "out is NULL");
    *out = nullptr;
    validate_params(build);
    if (max_jobs == 0) fail(GLX_EINVAL, "max_jobs must be > 0");
    ensure_device();
    glx_pool* p = new glx_pool();
    try {
      p->build = *build;
      p->budget = budget_bytes;
      p->max_jobs = max_jobs;
      p->st = make_stream();
      p->batch = new Batch(p->st);
      p->prime = new Batch(p->st);
      p->batch->set_mode(mode_from(p->build));
      p->prime->set_mode(mode_from(p->build));
      check(hipEventCreate(&p->ev0), "hipEventCreate");
      check(hipEventCreate(&p->ev1), "hipEventCreate");
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

extern "C" int glx_pool_destroy(glx_pool* p) {
  return guard([&] { delete p; });
}

// Admission shared by uploaded and generated jobs: checks, reserves, fills the state
// with load(desc) on the pool stream, and queues the job for priming.
template <class Load>
static void admit_job(glx_pool* p, size_t n, double g, double softening, double theta,
                      double dt, uint64_t steps, uint64_t* id, Load&& load) {
  glx_params jp = p->build;
  jp.g = g;
  jp.softening = softening;
  jp.theta = theta;
  validate_params(&jp);
  if (!isfinite(dt)) fail(GLX_EINVAL, "dt must be finite");
  if (p->jobs.size() >= p->max_jobs) fail(GLX_ENOSLOT, "no free job slot");
  const uint64_t need = Segment::bytes_for((uint32_t)n, false);
  if (p->used + need > p->budget) fail(GLX_EBUDGET, "device-memory budget exhausted");
  auto j = std::make_unique<Job>();
  j->seg.reserve((uint32_t)n, false);
  fill_params(j->seg.desc, jp, (uint32_t)n);
  j->seg.desc.dt = dt;
  j->seg.desc.half_dt = (float)(0.5 * dt);
  load(j->seg.desc);
  sync(p->st);  // a(x0) follows at the next refresh, batched with the other new jobs
  j->id = p->next_id++;
  j->activated = p->next_activation++;
  j->steps_target = steps;
  p->used += j->seg.bytes();
  *id = j->id;
  p->jobs.emplace(j->id, std::move(j));
  p->dirty = true;
}

extern "C" int glx_pool_admit(glx_pool* p, const glx_job* job, uint64_t* id) {
  return guard([&] {
    if (!p || !job || !id) fail(GLX_EINVAL, "NULL argument");
    *id = 0;
    validate_n(job->n);
    if (job->n == 0) fail(GLX_EINVAL, "a job needs at least one particle");
    if (!job->pos || !job->vel || !job->mass) fail(GLX_EINVAL, "NULL array");
    admit_job(p, job->n, job->g, job->softening, job->theta, job->dt, job->steps, id,
              [&](const SegDesc& d) {
                upload_state(d, job->n, job->pos, job->vel, job->mass, p->st);
              });
  });
}

extern "C" int glx_pool_admit_ic(glx_pool* p, const glx_ic_spec* spec, double softening,
                                 double theta, double dt, uint64_t steps, uint64_t* id) {
  return guard([&] {
    if (!p || !spec || !id) fail(GLX_EINVAL, "NULL argument");
    *id = 0;
    ic::validate_spec(*spec);
    admit_job(p, spec->n, ic::spec_g(*spec), softening, theta, dt, steps, id,
              [&](const SegDesc& d) { ic::generate_resident(*spec, d.x, d.v, d.m, p->st); });
  });
}

extern "C" int glx_pool_step(glx_pool* p, double* device_ms) {
  return guard([&] {
    if (!p) fail(GLX_EINVAL, "pool is NULL");
    if (device_ms) *device_ms = 0.0;
    pool_refresh(p);
    if (p->active.empty()) return;
    check(hipEventRecord(p->ev0, p->st), "hipEventRecord");
    p->graph.launch(p->st);
    check(hipEventRecord(p->ev1, p->st), "hipEventRecord");
    sync(p->st);
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, p->ev0, p->ev1), "hipEventElapsedTime");
    if (device_ms) *device_ms = ms;
    pool_advance(p, 1);
    pool_enforce_budget(p, ms);
  });
}

extern "C" int glx_pool_resume(glx_pool* p, uint64_t id) {
  return guard([&] {
    if (!p) fail(GLX_EINVAL, "pool is NULL");
    Job* j = find_job(p, id);
    if (!j->evicted) fail(GLX_EINVAL, "job is not evicted");
    j->evicted = false;  // its cached a(x) is that of its current state: no priming
    j->activated = p->next_activation++;
    p->dirty = true;
  });
}

extern "C" int glx_pool_set_step_budget(glx_pool* p, double budget_ms) {
  return guard([&] {
    if (!p) fail(GLX_EINVAL, "pool is NULL");
    if (!(budget_ms >= 0.0)) fail(GLX_EINVAL, "step budget must be >= 0 ms");
    p->step_budget_ms = budget_ms;
  });
}

extern "C" int glx_pool_step_timed(glx_pool* p, double* phase_ms) {
  return guard([&] {
    if (!p || !phase_ms) fail(GLX_EINVAL, "NULL argument");
    for (int q = 0; q < kPhases; ++q) phase_ms[q] = 0.0;
    pool_refresh(p);
    if (p->active.empty()) return;
    timed_steps(*p->batch, 1, phase_ms);
    pool_advance(p, 1);
  });
}

extern "C" int glx_pool_count(glx_pool* p, glx_stats* out) {
  return guard([&] {
    if (!p || !out) fail(GLX_EINVAL, "NULL argument");
    *out = glx_stats{};
    pool_refresh(p);
    if (p->active.empty()) return;
    for (Job* j : p->active) zero_walk_counters(j->seg.desc, p->st);
    Batch& b = *p->batch;
    b.reset_bbox();
    b.bbox();
    b.build_and_walk(0, true);
    for (Job* j : p->active) {
      unsigned long long c[kCounters] = {};
      d2h(c, j->seg.desc.counters, sizeof(c), p->st);
      sync(p->st);
      const glx_stats s = stats_from(c);
      out->point_terms += s.point_terms;
      out->quad_terms += s.quad_terms;
      out->node_tests += s.node_tests;
      out->fallbacks += s.fallbacks;
      out->max_stack = s.max_stack > out->max_stack ? s.max_stack : out->max_stack;
      out->far_nodes += s.far_nodes;
      out->rounds += s.rounds;
    }
  });
}

extern "C" int glx_pool_job(glx_pool* p, uint64_t id, glx_job_info* out) {
  return guard([&] {
    if (!p || !out) fail(GLX_EINVAL, "NULL argument");
    const Job* j = find_job(p, id);
    out->id = j->id;
    out->n = j->seg.desc.n;
    out->steps_done = j->steps_done;
    out->steps = j->steps_target;
    out->time = j->time;
    out->done = j->done ? 1 : 0;
    out->evicted = j->evicted ? 1 : 0;
  });
}

static void job_snapshot(glx_pool* p, const Job* j, double* pos, double* vel) {
  const size_t n = j->seg.desc.n;
  std::vector<float> v32(vel ? 3 * n : 0);
  if (pos) d2h(pos, j->seg.desc.x, 3 * n * sizeof(double), p->st);
  if (vel) d2h(v32.data(), j->seg.desc.v, 3 * n * sizeof(float), p->st);
  sync(p->st);
  if (vel)
    for (size_t q = 0; q < 3 * n; ++q) vel[q] = (double)v32[q];
}

extern "C" int glx_pool_snapshot(glx_pool* p, uint64_t id, double* pos, double* vel) {
  return guard([&] {
    if (!p) fail(GLX_EINVAL, "pool is NULL");
    job_snapshot(p, find_job(p, id), pos, vel);
  });
}

extern "C" int glx_pool_snapshot_accel(glx_pool* p, uint64_t id, double* acc) {
  return guard([&] {
    if (!p || !acc) fail(GLX_EINVAL, "NULL argument");
    const Job* j = find_job(p, id);
    pool_refresh(p);  // primes a job admitted since the last step
    const size_t n = j->seg.desc.n;
    std::vector<float> a32(3 * n);
    d2h(a32.data(), j->seg.desc.acc, 3 * n * sizeof(float), p->st);
    sync(p->st);
    for (size_t q = 0; q < 3 * n; ++q) acc[q] = (double)a32[q];
  });
}

extern "C" int glx_pool_snapshot_mass(glx_pool* p, uint64_t id, double* mass) {
  return guard([&] {
    if (!p || !mass) fail(GLX_EINVAL, "NULL argument");
    const Job* j = find_job(p, id);
    d2h(mass, j->seg.desc.m, (size_t)j->seg.desc.n * sizeof(double), p->st);
    sync(p->st);
  });
}

extern "C" int glx_pool_retire(glx_pool* p, uint64_t id, double* pos, double* vel) {
  return guard([&] {
    if (!p) fail(GLX_EINVAL, "pool is NULL");
    Job* j = find_job(p, id);
    job_snapshot(p, j, pos, vel);
    p->graph.drop();  // the graph references the job's memory
    p->used -= j->seg.bytes();
    p->jobs.erase(id);
    p->dirty = true;
  });
}

extern "C" int glx_pool_status(glx_pool* p, glx_pool_info* out) {
  return guard([&] {
    if (!p || !out) fail(GLX_EINVAL, "NULL argument");
    *out = glx_pool_info{};
    out->jobs = (uint32_t)p->jobs.size();
    for (auto& kv : p->jobs)
      if (!kv.second->done && !kv.second->evicted) {
        out->active += 1;
        out->particles += kv.second->seg.desc.n;
      }
    out->bytes_used = p->used;
    out->budget = p->budget;
    out->steps = p->steps;
    out->evictions = p->evictions;
    out->refresh_ms = p->refresh_ms;
    out->step_budget_ms = p->step_budget_ms;
  });
}

extern "C" uint64_t glx_pool_job_bytes(size_t n) {
  return n == 0 || n > kMaxN ? 0 : Segment::bytes_for((uint32_t)n, false);
}
