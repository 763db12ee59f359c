// GLXYSNAP snapshot streams (io/src/snapshot.rs), host side: writer and reader shared by
// the snapshot entry points and the stepping driver.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "galaxy_sim.h"

namespace glx::snap {

// A parsed stream; the columns are empty when only the header was read.
struct Parsed {
  uint32_t version = 0;
  double time = 0.0, scale_factor = 0.0, softening = 0.0;
  uint64_t step = 0, n = 0, rng_seed = 0, config_hash = 0;
  std::string units, code_version;
  std::vector<double> pos, vel, mass, u, formation_time;
  std::vector<uint64_t> id;
  std::vector<uint16_t> progenitor;
  std::vector<uint8_t> kind;
};

// Writes a version-4 stream of n particles (positions/velocities 3n, masses n; the
// optional columns default as documented in galaxy_sim.h). Throws glx::Error.
void write(const char* path, const glx_snap_header& h, size_t n, const double* pos,
           const double* vel, const double* mass, const glx_snap_columns* cols);

// Reads a version 1-4 stream. Throws glx::Error (GLX_EIO, GLX_EFORMAT).
Parsed read(const char* path, bool header_only);

}  // namespace glx::snap
