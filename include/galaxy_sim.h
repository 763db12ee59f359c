/*
 * galaxy_sim.h — the stepping driver and snapshot files around the LBVH step.
 *
 * glx_sys_run is sim::run (sim/src/lib.rs:110-174) over the resident stepper: emit the
 * state as uploaded, integrate n_steps KDK steps, emit a snapshot every snapshot_every
 * steps and always the final one. Snapshots are the reference's GLXYSNAP stream
 * (io/src/snapshot.rs:1-37): written at format version 4, read back from versions 1-4.
 * Host code only; the same status codes and glx_last_error() as galaxy_lbvh.h.
 */
#ifndef GALAXY_SIM_H
#define GALAXY_SIM_H

#include <stddef.h>
#include <stdint.h>

#include "galaxy_lbvh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GLX_EIO 7      /* a snapshot file could not be opened, read or written */
#define GLX_EFORMAT 8  /* not a snapshot stream, unsupported version, or corrupt */

/* io::Header (io/src/snapshot.rs:81-100). On write, n_particles is ignored (the count
 * comes from the data, io/src/snapshot.rs:130-160); units/code_version are NUL-terminated.
 * On read, units/code_version point into the caller's buffers. */
typedef struct glx_snap_header {
  double time;
  uint64_t step;
  double scale_factor;
  double softening;
  uint64_t n_particles;
  uint64_t rng_seed;
  uint64_t config_hash;
  const char *units;
  const char *code_version;
} glx_snap_header;

/* The State columns the gravity step does not evolve (core/src/state.rs:29-62); any
 * may be NULL: id = index, progenitor = 0, kind = 0 (Species::Collisionless),
 * u = 0.0, formation_time = -inf (State::PRIMORDIAL). */
typedef struct glx_snap_columns {
  const uint64_t *id;
  const uint16_t *progenitor;
  const uint8_t *kind;
  const double *u;
  const double *formation_time;
} glx_snap_columns;

/* io::write_file / to_writer (io/src/snapshot.rs:130-194,319-328): mass is stored as f32;
 * header strings are limited to 65536 bytes. */
int glx_snapshot_write(const char *path, const glx_snap_header *header, size_t n,
                       const double *pos, const double *vel, const double *mass,
                       const glx_snap_columns *columns);

/* The header of a snapshot (from_reader, io/src/snapshot.rs:199-221); version may be NULL.
 * A string longer than its buffer (NUL included) is GLX_EINVAL. */
int glx_snapshot_read_header(const char *path, glx_snap_header *header, char *units,
                             size_t units_len, char *code_version, size_t code_version_len,
                             uint32_t *version);

/* The columns of a snapshot whose header holds n particles (from_reader,
 * io/src/snapshot.rs:222-316, v1-v3 defaults included); any output may be NULL. */
int glx_snapshot_read(const char *path, size_t n, double *pos, double *vel, double *mass,
                      uint64_t *id, uint16_t *progenitor, uint8_t *kind, double *u,
                      double *formation_time);

/* sim::SimConfig (sim/src/lib.rs:36-59) without the SPH/star-formation fields. */
typedef struct glx_run_config {
  double dt;               /* > 0, finite */
  uint64_t n_steps;
  uint64_t snapshot_every; /* >= 1 */
  double softening;        /* recorded in headers */
  uint64_t rng_seed;       /* recorded in headers */
  uint64_t config_hash;    /* recorded in headers */
  const char *units;       /* recorded in headers (NULL = "") */
  uint64_t first_step;     /* step index of the uploaded state: 0 for sim::run, the
                              snapshot's step when resuming from one */
} glx_run_config;

/* sim::RunSummary (sim/src/lib.rs:62-70). */
typedef struct glx_run_summary {
  uint64_t steps;
  double final_time;
  uint64_t snapshots_emitted;
} glx_run_summary;

/* sim::run (sim/src/lib.rs:110-174) with a DirectorySink (sim/src/lib.rs:80-106): files
 * dir/snapshot_<step, 8 digits>.snap; header time from the stepper's clock. */
int glx_sys_run(glx_sys *sys, const glx_run_config *config, const char *dir,
                const glx_snap_columns *columns, glx_run_summary *out);

/* Sets the stepper's clock (State::time) — for resuming from a snapshot. */
int glx_sys_set_time(glx_sys *sys, double time);

#ifdef __cplusplus
}
#endif

#endif /* GALAXY_SIM_H */
