/*
 * galaxy_lbvh.h — C ABI of the MI355X-native LBVH gravity step.
 *
 * Drop-in boundary for the reference's LBVH path (BoykoNeov/galaxy). Each entry point
 * names the reference interface it replaces (file:line in the reference tree). Plain
 * pointers and sizes only; no C++ exceptions cross this boundary; every call returns
 * a status code and, on failure, leaves a message readable by glx_last_error().
 *
 * Layouts: positions, velocities and accelerations are interleaved xyz f64 arrays of
 * 3n doubles (the memory layout of a Rust &[DVec3]); masses are n doubles.
 *
 * Threading: a handle is used by one thread at a time (the reference's &mut self);
 * distinct handles may be used from distinct threads — each owns its own HIP stream.
 */
#ifndef GALAXY_LBVH_H
#define GALAXY_LBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GLX_OK 0
#define GLX_EINVAL 1    /* invalid argument (precondition the reference asserts) */
#define GLX_ENOMEM 2    /* device or host allocation failed */
#define GLX_EDEVICE 3   /* no usable GPU, or a HIP call failed */
#define GLX_EBUDGET 4   /* pool admission: device-memory budget exhausted */
#define GLX_ENOSLOT 5   /* pool admission: no free job slot */
#define GLX_ENOTFOUND 6 /* unknown job id */

/* walk and drift precision */
#define GLX_PREC_F32 0   /* f32 walk relative to the root-cube centre, plain accumulation;
                            f64 drift */
#define GLX_PREC_KAHAN 1 /* the same with Kahan-compensated accumulation */
#define GLX_PREC_ABS32 2 /* the reference's plain f32 path: f32 coordinates relative to
                            the coordinate origin, plain accumulation, f32 drift
                            (gpu/src/gpu_resident.rs:94-159 without double-single) */

/* flags */
#define GLX_FLAG_COUNT 1u /* count walk interactions / node visits (glx_solver_stats) */

/* walk */
#define GLX_WALK_REFERENCE 0 /* per-target walk, the reference acceptance test */
#define GLX_WALK_GROUP 1     /* wave-cooperative group walk (conservative group test) */

/* Solver parameters: the (g, softening, theta) of Lbvh::new
 * (solvers/src/lbvh.rs:652-659) plus this implementation's knobs. */
typedef struct glx_params {
  double g;             /* gravitational constant */
  double softening;     /* Plummer softening length eps (> = 0) */
  double theta;         /* opening angle (> 0) */
  uint32_t leaf_size;   /* bodies per leaf, 0 or 1 = the reference's LBVH leaf; up to 32
                           with GLX_WALK_GROUP: a node of at most leaf_size bodies that
                           fails the opening test is summed body by body */
  uint32_t multipole;   /* 0 = monopole (the reference's expansion), 2 = + quadrupole */
  uint32_t precision;   /* GLX_PREC_* */
  uint32_t flags;       /* GLX_FLAG_* */
  uint32_t morton_bits; /* 30 (reference, also 0) or 63 (2^21 cells per axis) */
  uint32_t walk;        /* GLX_WALK_*; quadrupoles need GLX_WALK_GROUP */
  double far_theta;     /* 0 = off; in (0, 1) with GLX_WALK_GROUP and multipole 2: an
                           accepted node whose radius plus the group's radius is at most
                           far_theta times its distance from the group centre enters one
                           second-order local expansion per group instead of a term per
                           body */
  double mono_ratio;    /* 0 = off; in (0, 1) with multipole 2: an accepted node that also
                           passes the opening test at theta * mono_ratio (distance >=
                           s / (theta mono_ratio) + delta from every point of the group) is
                           summed as a monopole alone; nearer accepted nodes keep the
                           quadrupole */
  /* Tuning (no reference counterpart; results do not depend on radix_bits): */
  uint32_t radix_bits;  /* sort digit width, 0 = 8; 8, 9 or 10 bits per LSD pass */
  uint32_t group_size;  /* GLX_WALK_GROUP block shape: bodies per walk group (the subtrees
                           of at most this many bodies), 0 = 64; 8..64 */
  uint32_t group_lanes; /* GLX_WALK_GROUP block shape: most lanes per body in groups of at
                           most 32 bodies, 0 = 4; 1..4 */
} glx_params;

/* FP64 conservation diagnostics (core/src/diagnostics.rs:8-68). */
typedef struct glx_diag {
  double kinetic;            /* 0.5 sum m v^2 */
  double potential;          /* softened potential (if requested), else 0 */
  double momentum[3];        /* sum m v */
  double angular_momentum[3];/* sum m r x v */
  double center_of_mass[3];  /* sum m r / sum m */
  double mass;               /* sum m */
} glx_diag;

/* Walk statistics (measurement hook; no reference counterpart). Terms are summed over
 * targets: one point-mass term is one leaf body or monopole node acting on one target
 * (the reference's interaction, solvers/src/lbvh.rs:490-517). */
typedef struct glx_stats {
  uint64_t point_terms; /* leaf-body and monopole-node terms */
  uint64_t quad_terms;  /* quadrupole-node terms */
  uint64_t node_tests;  /* nodes tested (per-target walk: per target; group walk: per group) */
  uint64_t fallbacks;   /* group walks that fell back to the per-target walk */
  uint64_t max_stack;   /* deepest group-walk node stack, entries */
  uint64_t far_nodes;   /* nodes folded into group local expansions (per group) */
  uint64_t rounds;      /* group-walk rounds (up to 64 node tests each), summed over groups */
} glx_stats;

/* Message of the last failed call on this thread ("" if none). */
const char *glx_last_error(void);

/* Device name (the gfx target where the driver reports no name), compute-unit count and
 * HBM bytes of device 0 (GLX_EDEVICE if none). */
int glx_device_info(char *name, size_t name_len, int *compute_units, uint64_t *hbm_bytes);

/* Device memory for callers without an allocator of their own (the Python front end uses
 * them for glx_ic_generate_device): hipMalloc / hipFree / a synchronous device-to-host
 * copy. No reference counterpart. */
int glx_device_malloc(size_t bytes, void **ptr);
int glx_device_free(void *ptr);
int glx_device_to_host(void *dst, const void *src, size_t bytes);

/* ---------------------------------------------------------------------------------
 * ForceSolver (core/src/traits.rs:24-27,52-54,95) as implemented by Lbvh
 * (solvers/src/lbvh.rs:643-683) and GpuLbvhFused (gpu/src/gpu_lbvh_fused.rs:50-134).
 * ------------------------------------------------------------------------------- */
typedef struct glx_solver glx_solver;

/* Lbvh::new / GpuLbvhFused::new (solvers/src/lbvh.rs:653; gpu/src/gpu_lbvh_fused.rs:50).
 * GLX_EINVAL for theta <= 0, softening < 0, non-finite g/eps/theta, leaf_size > 32 (or
 * > 1 without GLX_WALK_GROUP), unknown multipole/precision/morton_bits/walk;
 * GLX_EDEVICE without a GPU. */
int glx_solver_create(const glx_params *params, glx_solver **out);
int glx_solver_destroy(glx_solver *solver);

/* ForceSolver::accelerations (core/src/traits.rs:27; solvers/src/lbvh.rs:663-676;
 * gpu/src/gpu_lbvh_fused.rs:57-134). acc receives 3n doubles. n == 0 is a no-op and
 * n == 1 yields zero acceleration (gpu/src/gpu_lbvh.rs:410-418). Device buffers grow
 * on demand and never shrink (gpu/src/fused_core.rs:432). */
int glx_solver_accelerations(glx_solver *solver, size_t n, const double *pos,
                             const double *mass, double *acc);

/* ForceSolver::accelerations_active (core/src/traits.rs:43-54): a fresh tree of all n
 * bodies, walked only for the groups that hold an active body, so each active entry
 * equals the full pass bit for bit; entries outside `active` are unspecified (the
 * trait's contract). active == NULL walks every body. Index >= n: GLX_EINVAL. */
int glx_solver_accelerations_active(glx_solver *solver, size_t n, const double *pos,
                                    const double *mass, const size_t *active,
                                    size_t n_active, double *acc);

/* ForceSolver::rebuild_gravity_cache (core/src/traits.rs:73-78): builds the tree of these
 * positions and keeps it (every accelerations call also leaves its tree cached). */
int glx_solver_rebuild_cache(glx_solver *solver, size_t n, const double *pos,
                             const double *mass);

/* ForceSolver::gravity_active_cached (core/src/traits.rs:80-91): walks the kept tree for
 * the active bodies at the current positions — bodies of opened nodes at `pos`, accepted
 * nodes with the multipoles and bounds of the kept build. Entries outside `active` are
 * unspecified; GLX_EINVAL when no tree of n bodies is kept. */
int glx_solver_gravity_active_cached(glx_solver *solver, size_t n, const double *pos,
                                     const double *mass, const size_t *active,
                                     size_t n_active, double *acc);

/* ForceSolver::potential_energy (core/src/traits.rs:95; solvers/src/lbvh.rs:680-682):
 * exact O(n^2) FP64 softened potential, computed on the GPU. */
int glx_solver_potential_energy(glx_solver *solver, size_t n, const double *pos,
                                const double *mass, double *out);

/* Walk statistics of the last accelerations call (needs GLX_FLAG_COUNT). */
int glx_solver_stats(glx_solver *solver, glx_stats *out);

/* Exact FP64 direct sum (solvers/src/direct_sum.rs:25-45 pair kernel) onto a subset
 * of targets against all n sources, self excluded; acc_out holds 3 * n_targets. The
 * FP64 error reference for sampled targets at sizes where O(n^2) on the CPU is too
 * slow. */
int glx_direct_sum_f64(size_t n, const double *pos, const double *mass, double g,
                       double softening, const uint32_t *targets, size_t n_targets,
                       double *acc_out);

/* glx_direct_sum_f64 without any device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight serving loops that check
 * accuracy between steps, harnesses that forbid hipMalloc). Same arguments and the same
 * results, bit for bit. Its device memory (staged inputs, output, partial sums) is one
 * static 64 MiB array per GPU, held until the process ends; GLX_EINVAL when
 * 32 n + 28 n_targets bytes (each array rounded up to 256 B) exceed its 48 MiB staging
 * part. Calls on one GPU share the array: they are serialised inside the library. */
int glx_direct_sum_f64_noalloc(size_t n, const double *pos, const double *mass, double g,
                               double softening, const uint32_t *targets, size_t n_targets,
                               double *acc_out);

/* ---------------------------------------------------------------------------------
 * Stage exports: the GPU counterparts of the reference's stage oracles, which it makes
 * public for exactly this gating (solvers/src/lib.rs:15-18).
 * ------------------------------------------------------------------------------- */

/* reference_morton (solvers/src/lbvh.rs:105-129): bounds_out = {bmin xyz, size, scale};
 * lanes (3n, may be NULL) and codes (n). Bit-exact with the reference (f64). n > 0. */
int glx_stage_morton(size_t n, const double *pos, double *bounds_out, uint32_t *lanes,
                     uint32_t *codes);
/* 63-bit variant (2^21 cells per axis): the reference's deferred refinement
 * (solvers/src/lbvh.rs:29-31). */
int glx_stage_morton63(size_t n, const double *pos, double *bounds_out, uint32_t *lanes,
                       uint64_t *codes);

/* reference_sort (solvers/src/lbvh.rs:143-147): order sorting by (code, index), and
 * the sorted codes. Codes must be < 2^32 (30-bit Morton codes in practice). */
int glx_stage_sort(size_t n, const uint32_t *codes, uint32_t *order, uint32_t *sorted_codes);
int glx_stage_sort63(size_t n, const uint64_t *codes, uint32_t *order, uint64_t *sorted_codes);
/* The same with radix_bits-wide digits (8, 9 or 10; no reference counterpart): every
 * code must fit the passes' bits (32, 36, 30 for 32-bit codes; 64, 63, 70 for 64-bit). */
int glx_stage_sort_bits(size_t n, const uint32_t *codes, uint32_t *order,
                        uint32_t *sorted_codes, uint32_t radix_bits);
int glx_stage_sort63_bits(size_t n, const uint64_t *codes, uint32_t *order,
                          uint64_t *sorted_codes, uint32_t radix_bits);

/* reference_karras (solvers/src/lbvh.rs:276-306): left/right (n-1) and parent
 * (2n-1, unified index; parent[root] = 0xFFFFFFFF). n > 0. */
int glx_stage_karras(size_t n, const uint32_t *sorted_codes, uint32_t *left, uint32_t *right,
                     uint32_t *parent);
int glx_stage_karras63(size_t n, const uint64_t *sorted_codes, uint32_t *left, uint32_t *right,
                       uint32_t *parent);

/* reference_flatten (solvers/src/lbvh.rs:545-614; the walks here do not need it): the
 * tree of the sorted inputs laid out in DFS pre-order, left child first, 2n-1 slots:
 * next (the slot after the node's subtree), body_start / body_count (a leaf: its sorted
 * index and 1, so leaf_bodies[k] = order[k]; an internal node: 0 and 0), and f32
 * geometry relative to origin_out: center and half (3 per slot) of the AABB, com (3),
 * mass and delta = |com - center|. The structure is the reference's exactly. n > 0. */
int glx_stage_flatten(size_t n, const uint32_t *sorted_codes, const double *sorted_pos,
                      const double *sorted_mass, double *origin_out, uint32_t *next,
                      uint32_t *body_start, uint32_t *body_count, float *center, float *half,
                      float *com, float *mass, float *delta);

/* reference_aggregate (solvers/src/lbvh.rs:329-379) as computed by the device: per
 * unified node (2n-1) AABB min/max, centre of mass (3 floats each) and mass, in f32
 * coordinates relative to origin_out (the root cube centre of the sorted positions,
 * f64). quad (may be NULL) receives 6 floats per node, the traceless quadrupole
 * {xx, yy, zz, xy, xz, yz} = sum m (3 d d^T - |d|^2 I) about the node's centre of
 * mass (the reference is monopole-only, solvers/src/barnes_hut.rs:33). n > 0. */
int glx_stage_aggregate(size_t n, const uint32_t *sorted_codes, const double *sorted_pos,
                        const double *sorted_mass, double *origin_out, float *aabb_min,
                        float *aabb_max, float *com, float *mass, float *quad);

/* ---------------------------------------------------------------------------------
 * Resident stepper: GpuResidentLeapfrog (gpu/src/gpu_resident.rs:1246-1968). State
 * stays on the device between upload and snapshot; one KDK step is one hipGraph
 * launch. Positions are FP64 on the device, velocities and forces FP32.
 * ------------------------------------------------------------------------------- */
typedef struct glx_sys glx_sys;
struct glx_ic_spec; /* galaxy_ic.h */

/* GpuResidentLeapfrog::new (gpu/src/gpu_resident.rs:1292). */
int glx_sys_create(const glx_params *params, glx_sys **out);
int glx_sys_destroy(glx_sys *sys);
/* upload (gpu/src/gpu_resident.rs:1685): copies state, primes a(x0), resets time. */
int glx_sys_upload(glx_sys *sys, size_t n, const double *pos, const double *vel,
                   const double *mass);
/* As glx_sys_upload, with the state drawn on the device by a generator of galaxy_ic.h
 * (same values as glx_ic_generate on the host): no host arrays, no upload. The system
 * keeps the g it was created with. The ic crate samples on the host only
 * (ic/src/lib.rs:9-33); this is the device counterpart of upload(sample(...)). */
int glx_sys_generate(glx_sys *sys, const struct glx_ic_spec *spec);
/* step / step_many (gpu/src/gpu_resident.rs:1823,1848): k KDK steps of dt. */
int glx_sys_step(glx_sys *sys, double dt, uint64_t k);
/* As glx_sys_step, launched without the graph and with a HIP event between phases on
 * the system's stream; phase_ms[7] accumulates the device time of {kick-drift+bbox,
 * encode, sort, gather, Karras, aggregate, walk+kick}. The reference times only the
 * host side of step_many (gpu/tests/gpu_resident.rs:484-544); this is its device-side
 * measurement hook. */
int glx_sys_step_timed(glx_sys *sys, double dt, uint64_t k, double *phase_ms);
/* Rebuild at the current positions and count walk interactions / node visits
 * (measurement hook; no reference counterpart). Leaves the cached a(x) unchanged. */
int glx_sys_count(glx_sys *sys, glx_stats *out);
/* snapshot (gpu/src/gpu_resident.rs:1878): pos/vel (3n each, either may be NULL). */
int glx_sys_snapshot(glx_sys *sys, double *pos, double *vel);
/* snapshot_accel (gpu/src/gpu_resident.rs:1655): cached a (3n doubles). */
int glx_sys_snapshot_accel(glx_sys *sys, double *acc);
/* Masses (n doubles). The reference tracks them on the host (snapshot,
 * gpu/src/gpu_resident.rs:1876); here they are read from the device, where
 * glx_sys_generate writes them. */
int glx_sys_snapshot_mass(glx_sys *sys, double *mass);
/* time / len (gpu/src/gpu_resident.rs:1947,1960). */
int glx_sys_time(glx_sys *sys, double *time);
int glx_sys_len(glx_sys *sys, size_t *n);
/* FP64 device diagnostics; with_potential adds the exact O(n^2) FP64 potential. */
int glx_sys_diag(glx_sys *sys, int with_potential, glx_diag *out);

/* ---------------------------------------------------------------------------------
 * Batched and served forms. No reference counterpart: the reference steps one system
 * per GpuResidentLeapfrog (gpu/src/gpu_resident.rs:1848, up to 64 steps per submit).
 * A pool keeps many systems ("jobs") resident and advances every active job one KDK
 * step per glx_pool_step with one hipGraph launch of segment-aware kernels. Jobs are
 * admitted by a device-memory budget and a slot count and join/leave between steps
 * (continuous batching). A job's results do not depend on which other jobs share the
 * launch (every kernel works segment-locally).
 * ------------------------------------------------------------------------------- */
typedef struct glx_pool glx_pool;

typedef struct glx_job {
  size_t n;              /* particles (> 0) */
  const double *pos;     /* 3n */
  const double *vel;     /* 3n */
  const double *mass;    /* n */
  double g, softening, theta, dt;
  uint64_t steps;        /* KDK steps to run; 0 = run until retired */
} glx_job;

typedef struct glx_job_info {
  uint64_t id;
  size_t n;
  uint64_t steps_done, steps;
  double time;
  int done;              /* steps_done reached steps: no longer advanced */
  int evicted;           /* paused by the pool's step budget (glx_pool_set_step_budget):
                            not advanced until glx_pool_resume; state kept until retired */
} glx_job_info;

typedef struct glx_pool_info {
  uint32_t jobs, active;
  uint64_t particles;    /* in active jobs */
  uint64_t bytes_used, budget;
  uint64_t steps;        /* pool steps taken */
  uint64_t evictions;    /* evictions by the step budget so far */
  double refresh_ms;     /* host time of the last step's refresh (a(x0) of newly admitted
                            jobs, segment tables and graph capture after the active set
                            changed; 0 when it did not change) */
  double step_budget_ms; /* glx_pool_set_step_budget (0: none) */
} glx_pool_info;

/* build: walk/multipole/morton_bits/precision shared by every job (g, softening and
 * theta are taken per job). */
int glx_pool_create(const glx_params *build, uint64_t budget_bytes, uint32_t max_jobs,
                    glx_pool **out);
int glx_pool_destroy(glx_pool *pool);
/* Uploads the job; its a(x0) is computed before the next step, in one batched launch with
 * the other jobs admitted since. GLX_EBUDGET / GLX_ENOSLOT when it does not fit. */
int glx_pool_admit(glx_pool *pool, const glx_job *job, uint64_t *id);
/* As glx_pool_admit, with the job's state drawn on the device by a generator of
 * galaxy_ic.h (no host arrays, no upload); the job's g is the generator's (spec->g, 1 for
 * a collision). */
int glx_pool_admit_ic(glx_pool *pool, const struct glx_ic_spec *spec, double softening,
                      double theta, double dt, uint64_t steps, uint64_t *id);
/* One KDK step of every active job; device_ms (may be NULL) = device time of the step. */
int glx_pool_step(glx_pool *pool, double *device_ms);
/* Per-step time budget (device ms of one glx_pool_step; 0 = none, the default). After a
 * step that took longer, the most recently admitted (or resumed) active job is evicted:
 * paused — no longer advanced, glx_pool_job reports evicted = 1 — with its state and
 * memory kept until glx_pool_resume or glx_pool_retire. One job per over-budget step, so
 * the pool sheds load until its steps fit. The reference bounds device work per submit for
 * the same reason (MAX_BATCH, gpu/src/gpu_resident.rs:1271-1287); a server keeps steps
 * within the budget by admitting jobs by their estimated cost (galaxy_amd/serve.py), and
 * eviction is the guard for what the estimate misses. GLX_EINVAL for a negative or NaN
 * budget. */
int glx_pool_set_step_budget(glx_pool *pool, double budget_ms);
/* Resume an evicted job: it is advanced again from its kept state (its cached a(x) is
 * that state's) with the next step. GLX_EINVAL if the job is not evicted. */
int glx_pool_resume(glx_pool *pool, uint64_t id);
/* One step with direct launches and per-phase device ms (see glx_sys_step_timed). */
int glx_pool_step_timed(glx_pool *pool, double *phase_ms);
/* Walk statistics summed over active jobs at their current positions. */
int glx_pool_count(glx_pool *pool, glx_stats *out);
int glx_pool_job(glx_pool *pool, uint64_t id, glx_job_info *out);
int glx_pool_snapshot(glx_pool *pool, uint64_t id, double *pos, double *vel);
int glx_pool_snapshot_accel(glx_pool *pool, uint64_t id, double *acc);
/* Masses (n doubles), as glx_sys_snapshot_mass: a job drawn on the device has its masses
 * only there. */
int glx_pool_snapshot_mass(glx_pool *pool, uint64_t id, double *mass);
/* Copies the job out (pos/vel may be NULL) and frees its device memory. */
int glx_pool_retire(glx_pool *pool, uint64_t id, double *pos, double *vel);
int glx_pool_status(glx_pool *pool, glx_pool_info *out);
/* Device bytes a job of n particles takes against the budget (0 if n is invalid). */
uint64_t glx_pool_job_bytes(size_t n);

#ifdef __cplusplus
}
#endif
#endif /* GALAXY_LBVH_H */
