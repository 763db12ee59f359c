// Per-particle samplers of the workload generators, shared by the host entry points and
// the device kernels (glx_ic.hip). Every particle draws from its own SplitMix64 stream
// keyed by (seed, component, index), so a particle's state is a pure function of its
// index: the device computes it in one thread, twice if it must (moments, then the final
// write), and the host computes the same values.
//
// Samplers follow the reference's ic crate:
//   SplitMix64, random direction     ic/src/plummer.rs:143-174, ic/src/disk.rs:463-479
//   Plummer (Aarseth-Henon-Wielen)   ic/src/plummer.rs:94-139
//   Hernquist (DF rejection, 300a)   ic/src/hernquist.rs:78-96,124-176
//   exponential sech^2 disk          ic/src/disk.rs:270-290,333-371,418-431
//   Kepler placement                 ic/src/encounter.rs:45-87
//   orientation Rz(arg) * Rx(incl)   ic/src/orient.rs:63-71
//
// FP contraction is off for everything below, so host and device evaluate the same
// operation sequence; they differ only where the math library's last bit does.
#ifndef GLX_IC_SAMPLER_H
#define GLX_IC_SAMPLER_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace glx {
namespace ic {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTau = 2.0 * kPi;
constexpr int kChunk = 256;         // particles per moment partial (one block)
constexpr int kGuard = 1 << 16;     // bound on every rejection loop
constexpr int kMoments = 7;         // sum m, sum m x (3), sum m v (3)

enum Kind : int { kPlummer = 0, kHernquist = 1, kCollision = 2 };
enum Comp : uint64_t { kCompDisk = 0, kCompBulge = 1, kCompHalo = 2, kCompPlummer = 3 };

struct V3 {
  double x, y, z;
};

// ic/src/plummer.rs:153-174
struct SplitMix64 {
  uint64_t state;
  __host__ __device__ explicit SplitMix64(uint64_t s) : state(s) {}
  __host__ __device__ uint64_t next_u64() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  __host__ __device__ double next_f64() {
    return (double)(next_u64() >> 11) / 9007199254740992.0;
  }
};

// ic/src/disk.rs:474-479
__host__ __device__ inline uint64_t mix_seed(uint64_t seed) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t stream_seed(uint64_t seed, uint64_t comp, uint64_t i) {
  return mix_seed(mix_seed(seed ^ (0xA0761D6478BD642Full * (comp + 1))) ^
                  (i * 0xE7037ED1A0B428DBull));
}

// ic/src/plummer.rs:143-148
__host__ __device__ inline V3 random_direction(SplitMix64& r) {
  const double ct = 2.0 * r.next_f64() - 1.0;
  const double st = sqrt(fmax(1.0 - ct * ct, 0.0));
  const double ph = kTau * r.next_f64();
  return {st * cos(ph), st * sin(ph), ct};
}

// ic/src/disk.rs:463-467
__host__ __device__ inline void box_muller(double u1, double u2, double& g1, double& g2) {
  const double rad = sqrt(-2.0 * log(fmax(u1, 2.2250738585072014e-308)));
  g1 = rad * cos(kTau * u2);
  g2 = rad * sin(kTau * u2);
}

// Plummer sphere truncated at rmax: X = M(<r)/M uniform on [0, x_max).
struct Plummer {
  double g, m, a, x_max;
  // ic/src/plummer.rs:103-123
  __host__ __device__ void sample_one(SplitMix64& r, V3& p, V3& v) const {
    const double x = x_max * r.next_f64();
    const double rad = a / sqrt(pow(x, -2.0 / 3.0) - 1.0);
    const V3 d = random_direction(r);
    p = {d.x * rad, d.y * rad, d.z * rad};
    double q = 0.0;
    for (int guard = 0; guard < kGuard; ++guard) {
      const double qq = r.next_f64();
      const double h = 0.1 * r.next_f64();
      if (h < qq * qq * pow(1.0 - qq * qq, 3.5)) {
        q = qq;
        break;
      }
    }
    const double ve = sqrt(2.0 * g * m) * pow(rad * rad + a * a, -0.25);
    const V3 dv = random_direction(r);
    v = {dv.x * (q * ve), dv.y * (q * ve), dv.z * (q * ve)};
  }
};

// Hernquist sphere, isotropic DF (ic/src/hernquist.rs:78-96), truncated at 300a.
struct Hernquist {
  double g, m, a;
  __host__ __device__ double enclosed(double r) const {
    return m * r * r / ((r + a) * (r + a));
  }
  __host__ __device__ double df(double energy) const {
    if (energy <= 0.0) return 0.0;
    const double gm = g * m, vg2 = gm / a, q = energy / vg2;
    if (q >= 1.0) return 0.0;
    const double vg = sqrt(vg2);
    const double pref = m / (8.0 * sqrt(2.0) * kPi * kPi * kPi * a * a * a * vg * vg * vg);
    const double br = 3.0 * asin(sqrt(q)) +
                      sqrt(q * (1.0 - q)) * (1.0 - 2.0 * q) * (8.0 * q * q - 8.0 * q - 3.0);
    return pref * br / pow(1.0 - q, 2.5);
  }
  // ic/src/hernquist.rs:139-176
  __host__ __device__ void sample_one(SplitMix64& r, V3& p, V3& v) const {
    const double x_max = (300.0 / 301.0) * (300.0 / 301.0);
    const double x = x_max * r.next_f64();
    const double rx = sqrt(x);
    const double rad = a * rx / (1.0 - rx);
    const V3 d = random_direction(r);
    p = {d.x * rad, d.y * rad, d.z * rad};
    const double psi = g * m / (rad + a);
    const double vesc = sqrt(2.0 * psi);
    double ceil = 0.0;
    for (int k = 1; k < 256; ++k) {
      const double vv = vesc * k / 256.0;
      ceil = fmax(ceil, vv * vv * df(psi - 0.5 * vv * vv));
    }
    ceil *= 1.5;
    double sp = 0.0;
    for (int guard = 0; guard < kGuard; ++guard) {
      const double vv = vesc * r.next_f64();
      const double h = ceil * r.next_f64();
      if (h < vv * vv * df(psi - 0.5 * vv * vv)) {
        sp = vv;
        break;
      }
    }
    const V3 dv = random_direction(r);
    v = {dv.x * sp, dv.y * sp, dv.z * sp};
  }
};

// One collision galaxy: exponential sech^2 disk (M 1, R_d 3, h 0.3, cut at 4 R_d) in a
// Hernquist bulge (M 0.33, a 0.5) and halo (M 5, a 10), every mass times `scale`; disk
// velocity dispersion 10% of v_c in each direction. Particles: disk 25%, bulge 10%
// (rounded), halo the rest, in that index order.
struct Galaxy {
  double g, scale;
  double disk_mass, rd, hz, rmax, sigma_frac;
  Hernquist bulge, halo;
  uint64_t n, nd, nb;
  uint64_t seed;

  __host__ __device__ double trunc_frac() const {
    const double u = rmax / rd;
    return 1.0 - (1.0 + u) * exp(-u);
  }
  // ic/src/disk.rs:270-274
  __host__ __device__ double disk_enclosed(double r) const {
    const double x = fmin(r, rmax) / rd;
    return disk_mass * (1.0 - (1.0 + x) * exp(-x)) / trunc_frac();
  }
  // ic/src/disk.rs:282-288, with the bulge, halo and disk enclosed masses
  __host__ __device__ double vcirc(double r) const {
    if (r <= 0.0) return 0.0;
    const double menc = bulge.enclosed(r) + halo.enclosed(r) + disk_enclosed(r);
    return sqrt(g * menc / r);
  }
  // ic/src/disk.rs:418-431: invert M_d(<R) = x M_d by bisection
  __host__ __device__ double sample_radius(double x) const {
    double lo = 0.0, hi = rmax;
    const double target = x * disk_mass;
    for (int k = 0; k < 60; ++k) {
      const double mid = 0.5 * (lo + hi);
      if (disk_enclosed(mid) < target)
        lo = mid;
      else
        hi = mid;
    }
    return 0.5 * (lo + hi);
  }
  // Particle i of this galaxy in its body frame (before recentring).
  __host__ __device__ void sample(uint64_t i, V3& p, V3& v, double& mass) const {
    if (i < nd) {
      // ic/src/disk.rs:333-371 (radius, azimuth, sech^2 height; dispersion about v_c)
      SplitMix64 r(stream_seed(seed, kCompDisk, i));
      const double R = sample_radius(r.next_f64());
      const double phi = kTau * r.next_f64();
      const double y = r.next_f64();
      const double t = fmin(fmax(2.0 * y - 1.0, -1.0 + 1e-12), 1.0 - 1e-12);
      const double z = hz * atanh(t);
      const double cp = cos(phi), sp = sin(phi);
      p = {R * cp, R * sp, z};
      const double vc = vcirc(R);
      double gr, gz, gphi, gunused;
      box_muller(r.next_f64(), r.next_f64(), gr, gz);
      box_muller(r.next_f64(), r.next_f64(), gphi, gunused);
      const double sig = sigma_frac * vc;
      const double vr = sig * gr, vphi = vc + sig * gphi, vz = sig * gz;
      v = {vr * cp - vphi * sp, vr * sp + vphi * cp, vz};
      mass = disk_mass / (double)nd;
    } else if (i < nd + nb) {
      SplitMix64 r(stream_seed(seed, kCompBulge, i));
      bulge.sample_one(r, p, v);
      mass = bulge.m / (double)nb;
    } else {
      SplitMix64 r(stream_seed(seed, kCompHalo, i));
      halo.sample_one(r, p, v);
      mass = halo.m / (double)(n - nd - nb);
    }
  }
};

// Rz(arg) * Rx(incl) applied to v (ic/src/orient.rs:63-71)
__host__ __device__ inline V3 orient(V3 v, double incl, double arg) {
  const double ci = cos(incl), si = sin(incl), ca = cos(arg), sa = sin(arg);
  const V3 r1{v.x, ci * v.y - si * v.z, si * v.y + ci * v.z};
  return {ca * r1.x - sa * r1.y, sa * r1.x + ca * r1.y, r1.z};
}

// A generator run: up to two sub-systems ("parts") of consecutive particles, each
// recentred on its own mass-weighted centre, then rotated and offset.
struct Plan {
  int kind;
  int parts;
  uint64_t n;
  uint64_t off[2], cnt[2];   // particle range of each part
  uint64_t blk0[2], blks[2]; // moment blocks of each part (kChunk particles each)
  Plummer plummer;
  Hernquist hernquist;
  Galaxy galaxy[2];
  uint64_t seed;             // Plummer / Hernquist streams
  double incl[2], arg[2];    // collision orientations
  double rp, r0;             // collision orbit
};

// Final transform of one part: p' = orient(p - com) + dp, v' = orient(v - cov) + dv.
struct Frame {
  double com[3], cov[3], dp[3], dv[3];
  double incl, arg;
  int rotate;
};

// Particle k of part `part` (local index) before recentring.
__host__ __device__ inline void plan_sample(const Plan& pl, int part, uint64_t k, V3& p, V3& v,
                                            double& mass) {
  if (pl.kind == kCollision) {
    pl.galaxy[part].sample(k, p, v, mass);
  } else if (pl.kind == kPlummer) {
    SplitMix64 r(stream_seed(pl.seed, kCompPlummer, k));
    pl.plummer.sample_one(r, p, v);
    mass = pl.plummer.m / (double)pl.n;
  } else {
    SplitMix64 r(stream_seed(pl.seed, kCompHalo, k));
    pl.hernquist.sample_one(r, p, v);
    mass = pl.hernquist.m / (double)pl.n;
  }
}

// The seven moments of one particle.
__host__ __device__ inline void moments_of(const V3& p, const V3& v, double m,
                                           double out[kMoments]) {
  out[0] = m;
  out[1] = m * p.x;
  out[2] = m * p.y;
  out[3] = m * p.z;
  out[4] = m * v.x;
  out[5] = m * v.y;
  out[6] = m * v.z;
}

// Frames from the summed moments of each part (sums[part][kMoments]).
__host__ __device__ inline void make_frames(const Plan& pl, const double sums[2][kMoments],
                                            Frame fr[2]) {
  for (int q = 0; q < pl.parts; ++q) {
    const double mt = sums[q][0];
    for (int c = 0; c < 3; ++c) {
      fr[q].com[c] = sums[q][1 + c] / mt;
      fr[q].cov[c] = sums[q][4 + c] / mt;
      fr[q].dp[c] = 0.0;
      fr[q].dv[c] = 0.0;
    }
    fr[q].incl = 0.0;
    fr[q].arg = 0.0;
    fr[q].rotate = 0;
  }
  if (pl.kind != kCollision) return;
  // ic/src/encounter.rs:45-87 with e = 1 (parabolic), G = 1
  const double m1 = sums[0][0], m2 = sums[1][0];
  const double mu = m1 + m2, e = 1.0;
  const double p = pl.rp * (1.0 + e), h = sqrt(mu * p);
  const double cnu = fmin(fmax((p / pl.r0 - 1.0) / e, -1.0), 1.0);
  const double nu = -acos(cnu);
  const double sn = sin(nu), cn = cos(nu);
  const double vr = mu / h * e * sn, vt = mu / h * (1.0 + e * cn);
  const double rrel[3] = {cn * pl.r0, sn * pl.r0, 0.0};
  const double vrel[3] = {cn * vr - sn * vt, sn * vr + cn * vt, 0.0};
  const double f[2] = {-m2 / (m1 + m2), m1 / (m1 + m2)};
  for (int q = 0; q < 2; ++q) {
    for (int c = 0; c < 3; ++c) {
      fr[q].dp[c] = rrel[c] * f[q];
      fr[q].dv[c] = vrel[c] * f[q];
    }
    fr[q].incl = pl.incl[q];
    fr[q].arg = pl.arg[q];
    fr[q].rotate = 1;
  }
}

// Final state of one particle: recentred, oriented, offset, rounded to f32 values.
__host__ __device__ inline void apply_frame(const Frame& f, V3& p, V3& v) {
  V3 pp{p.x - f.com[0], p.y - f.com[1], p.z - f.com[2]};
  V3 vv{v.x - f.cov[0], v.y - f.cov[1], v.z - f.cov[2]};
  if (f.rotate) {
    pp = orient(pp, f.incl, f.arg);
    vv = orient(vv, f.incl, f.arg);
  }
  p = {(double)(float)(pp.x + f.dp[0]), (double)(float)(pp.y + f.dp[1]),
       (double)(float)(pp.z + f.dp[2])};
  v = {(double)(float)(vv.x + f.dv[0]), (double)(float)(vv.y + f.dv[1]),
       (double)(float)(vv.z + f.dv[2])};
}

}  // namespace ic
}  // namespace glx

#endif  // GLX_IC_SAMPLER_H
