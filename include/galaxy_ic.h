/*
 * galaxy_ic.h — initial-condition generators for the benchmark workloads.
 *
 * Counterparts of the reference's samplers (ic crate); they build inputs, they are not
 * part of the LBVH step. Output arrays are interleaved xyz f64 (3n) and masses (n);
 * positions and velocities are rounded to f32-representable values.
 *
 * glx_ic_plummer keeps the reference's sequential draw order. Every generator named by a
 * glx_ic_spec draws each particle from its own SplitMix64 stream keyed by (seed,
 * component, index), so the same spec gives the same system on the host
 * (glx_ic_generate) and on the device (glx_ic_generate_device, glx_pool_admit_ic,
 * glx_sys_generate), up to the last bit of the math library's transcendental functions.
 */
#ifndef GALAXY_IC_H
#define GALAXY_IC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Plummer::sample (ic/src/plummer.rs:94-139), same SplitMix64 draw order, with the
 * radius truncated at rmax (X drawn on [0, M(<rmax)/M), as ic/src/hernquist.rs:139-148
 * truncates); rmax <= 0 means untruncated. Recentred to zero COM and momentum.
 * Returns 0 on success, 1 on invalid arguments. */
int glx_ic_plummer(size_t n, uint64_t seed, double g, double m, double a, double rmax,
                   double *pos, double *vel, double *mass);

/* Two-galaxy collision: each galaxy an exponential sech^2 disk + Hernquist bulge +
 * Hernquist halo (ic/src/disk.rs:307-411, ic/src/hernquist.rs:124-192), placed on a
 * parabolic Kepler orbit (ic/src/encounter.rs:45-87) with seeded random orientations
 * (ic/src/orient.rs:63-71). out_info (may be NULL) receives {incl1, arg1, incl2, arg2}.
 * Equal to glx_ic_generate with GLX_IC_COLLISION. Returns 0 on success, 1 on invalid
 * arguments. */
int glx_ic_collision(size_t n, uint64_t seed, double mass_ratio, double rp, double r0,
                     double *pos, double *vel, double *mass, double *out_info);

#define GLX_IC_PLUMMER 0   /* Plummer::sample (ic/src/plummer.rs:94), truncated at rmax */
#define GLX_IC_HERNQUIST 1 /* Hernquist::sample (ic/src/hernquist.rs:124), cut at 300a */
#define GLX_IC_COLLISION 2 /* the two-galaxy collision of glx_ic_collision, G = 1 */

typedef struct glx_ic_spec {
  int kind;                  /* GLX_IC_* */
  size_t n;                  /* particles: >= 1 (>= 2 for a collision), <= 2^29 */
  uint64_t seed;
  double g, m, a;            /* Plummer / Hernquist: G, total mass, scale radius (> 0) */
  double rmax;               /* Plummer truncation radius (<= 0: none) */
  double mass_ratio, rp, r0; /* collision: mass scale of galaxy 2 (> 0), pericentre and
                                starting separation (0 < rp <= r0) */
} glx_ic_spec;

/* Host generation: each particle drawn from its own stream, each part (the sphere, or
 * each galaxy) recentred on its mass-weighted centre with moments summed in a fixed
 * order (blocks of 256 particles reduced by halving, then the blocks in order).
 * out_info (may be NULL): {incl1, arg1, incl2, arg2} of a collision, zeros otherwise.
 * GLX_OK, or GLX_EINVAL (galaxy_lbvh.h) with glx_last_error(). */
int glx_ic_generate(const glx_ic_spec *spec, double *pos, double *vel, double *mass,
                    double *out_info);

/* The same generator on the GPU, written to device memory: pos and vel 3n f64, mass n
 * f64 (device pointers, e.g. hipMalloc or a framework tensor). No host arrays and no
 * upload; returns when the arrays are written. GLX_EDEVICE without a GPU. The ic crate
 * samples on the host only; this is its device counterpart. */
int glx_ic_generate_device(const glx_ic_spec *spec, double *pos, double *vel, double *mass);

#ifdef __cplusplus
}
#endif
#endif /* GALAXY_IC_H */
