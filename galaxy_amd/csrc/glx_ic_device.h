// Device generation of a system straight into resident arrays (glx_ic.hip), for the
// pool and the resident stepper.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "galaxy_ic.h"

namespace glx {
namespace ic {

// Draw spec's particles into device arrays laid out as the resident state: pos 3n f64,
// vel 3n f32, mass n f64; on stream st, synchronized before returning. Invalid specs
// raise GLX_EINVAL.
void generate_resident(const glx_ic_spec& s, double* pos, float* vel, double* mass,
                       hipStream_t st);
// Validate spec without drawing (GLX_EINVAL on failure).
void validate_spec(const glx_ic_spec& s);
// Gravitational constant of the generated system.
double spec_g(const glx_ic_spec& s);

}  // namespace ic
}  // namespace glx
