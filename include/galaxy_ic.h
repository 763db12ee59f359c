/*
 * galaxy_ic.h — initial-condition generators for the benchmark workloads.
 *
 * Counterparts of the reference's samplers (ic crate); they build inputs, they are not
 * part of the LBVH step. Output arrays are interleaved xyz f64 (3n) and masses (n);
 * positions and velocities are rounded to f32-representable values. Return 0 on
 * success, 1 on invalid arguments.
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
 * truncates); rmax <= 0 means untruncated. Recentred to zero COM and momentum. */
int glx_ic_plummer(size_t n, uint64_t seed, double g, double m, double a, double rmax,
                   double *pos, double *vel, double *mass);

/* Two-galaxy collision: each galaxy an exponential sech^2 disk + Hernquist bulge +
 * Hernquist halo (ic/src/disk.rs:307-411, ic/src/hernquist.rs:124-192), placed on a
 * parabolic Kepler orbit (ic/src/encounter.rs:45-87) with seeded random orientations
 * (ic/src/orient.rs:63-71). out_info (may be NULL) receives {incl1, arg1, incl2, arg2}. */
int glx_ic_collision(size_t n, uint64_t seed, double mass_ratio, double rp, double r0,
                     double *pos, double *vel, double *mass, double *out_info);

#ifdef __cplusplus
}
#endif
#endif /* GALAXY_IC_H */
