// Initial-condition generators for the benchmark workloads (host C++).
//
// Samplers follow the reference's ic crate:
//   Plummer (Aarseth-Henon-Wielen)   ic/src/plummer.rs:94-174
//   Hernquist (DF rejection)         ic/src/hernquist.rs:78-96,124-192
//   exponential sech^2 disk          ic/src/disk.rs:270-290,307-437,463-479
//   Kepler placement                 ic/src/encounter.rs:45-87
//   orientation Rz(arg) * Rx(incl)   ic/src/orient.rs:63-71
// The Plummer sampler reproduces the reference's sequential SplitMix64 draw order.
// The galaxy components draw from one SplitMix64 stream per particle (keyed by seed,
// component and index) so that millions of particles sample in parallel threads.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "galaxy_ic.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTau = 2.0 * kPi;

// ic/src/plummer.rs:153-174
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t s) : state(s) {}
  uint64_t next_u64() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double next_f64() { return (double)(next_u64() >> 11) / 9007199254740992.0; }
};

// ic/src/disk.rs:474-479
uint64_t mix_seed(uint64_t seed) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct V3 {
  double x, y, z;
};

// ic/src/plummer.rs:143-148
V3 random_direction(SplitMix64& r) {
  const double ct = 2.0 * r.next_f64() - 1.0;
  const double st = sqrt(std::max(1.0 - ct * ct, 0.0));
  const double ph = kTau * r.next_f64();
  return {st * cos(ph), st * sin(ph), ct};
}

// ic/src/disk.rs:463-467
void box_muller(double u1, double u2, double& g1, double& g2) {
  const double rad = sqrt(-2.0 * log(std::max(u1, 2.2250738585072014e-308)));
  g1 = rad * cos(kTau * u2);
  g2 = rad * sin(kTau * u2);
}

template <class F>
void parallel_for(size_t n, F&& f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  nt = std::min<unsigned>(nt, 16);
  if (n < 4096 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& x : th) x.join();
}

// Hernquist isotropic DF (ic/src/hernquist.rs:78-96).
struct Hernquist {
  double g, m, a;
  double enclosed(double r) const { return m * r * r / ((r + a) * (r + a)); }
  double df(double energy) const {
    if (energy <= 0.0) return 0.0;
    const double gm = g * m, vg2 = gm / a, q = energy / vg2;
    if (q >= 1.0) return 0.0;
    const double vg = sqrt(vg2);
    const double pref = m / (8.0 * sqrt(2.0) * kPi * kPi * kPi * a * a * a * vg * vg * vg);
    const double br = 3.0 * asin(sqrt(q)) +
                      sqrt(q * (1.0 - q)) * (1.0 - 2.0 * q) * (8.0 * q * q - 8.0 * q - 3.0);
    return pref * br / pow(1.0 - q, 2.5);
  }
  // ic/src/hernquist.rs:139-176, truncated at 300a
  void sample_one(SplitMix64& r, V3& p, V3& v) const {
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
      ceil = std::max(ceil, vv * vv * df(psi - 0.5 * vv * vv));
    }
    ceil *= 1.5;
    double sp = 0.0;
    for (int guard = 0; guard < 100000; ++guard) {
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

void narrow_f32(double* a, size_t n) {
  for (size_t i = 0; i < n; ++i) a[i] = (double)(float)a[i];
}

}  // namespace

// ic/src/plummer.rs:94-139 with truncation at rmax (X uniform on [0, M(<rmax)/M)).
extern "C" int glx_ic_plummer(size_t n, uint64_t seed, double g, double m, double a, double rmax,
                              double* pos, double* vel, double* mass) {
  if (n == 0) return 0;
  if (!pos || !vel || !mass || !(g > 0) || !(m > 0) || !(a > 0)) return 1;
  SplitMix64 rng(seed);
  const double x_max = (rmax > 0 && isfinite(rmax))
                           ? rmax * rmax * rmax / pow(rmax * rmax + a * a, 1.5)
                           : 1.0;
  const double vcoef = sqrt(2.0 * g * m);
  for (size_t i = 0; i < n; ++i) {
    const double x = x_max * rng.next_f64();
    const double r = a / sqrt(pow(x, -2.0 / 3.0) - 1.0);
    const V3 d = random_direction(rng);
    pos[3 * i] = d.x * r;
    pos[3 * i + 1] = d.y * r;
    pos[3 * i + 2] = d.z * r;
    double q = 0.0;
    for (;;) {
      const double qq = rng.next_f64();
      const double h = 0.1 * rng.next_f64();
      if (h < qq * qq * pow(1.0 - qq * qq, 3.5)) {
        q = qq;
        break;
      }
    }
    const double ve = vcoef * pow(r * r + a * a, -0.25);
    const V3 dv = random_direction(rng);
    vel[3 * i] = dv.x * (q * ve);
    vel[3 * i + 1] = dv.y * (q * ve);
    vel[3 * i + 2] = dv.z * (q * ve);
  }
  const double inv_n = 1.0 / (double)n;
  double mp[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      mp[c] += pos[3 * i + c];
      mv[c] += vel[3 * i + c];
    }
  for (int c = 0; c < 3; ++c) {
    mp[c] *= inv_n;
    mv[c] *= inv_n;
  }
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] -= mp[c];
      vel[3 * i + c] -= mv[c];
    }
  const double me = m / (double)n;
  for (size_t i = 0; i < n; ++i) mass[i] = me;
  narrow_f32(pos, 3 * n);
  narrow_f32(vel, 3 * n);
  return 0;
}

namespace {

struct GalaxyModel {
  double g = 1.0;
  double disk_mass = 1.0, rd = 3.0, hz = 0.3, rmax = 12.0;  // r_max = 4 R_d
  Hernquist bulge{1.0, 0.33, 0.5};
  Hernquist halo{1.0, 5.0, 10.0};
  double sigma_frac = 0.1;
  double scale = 1.0;  // component mass multiplier

  double trunc_frac() const {
    const double u = rmax / rd;
    return 1.0 - (1.0 + u) * exp(-u);
  }
  // ic/src/disk.rs:270-274
  double disk_enclosed(double r) const {
    const double x = std::min(r, rmax) / rd;
    return disk_mass * (1.0 - (1.0 + x) * exp(-x)) / trunc_frac();
  }
  // ic/src/disk.rs:282-288 with bulge + halo + disk enclosed mass
  double vcirc(double r) const {
    if (r <= 0.0) return 0.0;
    const double menc = bulge.enclosed(r) + halo.enclosed(r) + disk_enclosed(r);
    return sqrt(g * menc / r);
  }
  // ic/src/disk.rs:418-431
  double sample_radius(double x) const {
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
};

uint64_t stream_seed(uint64_t seed, uint64_t comp, uint64_t i) {
  return mix_seed(mix_seed(seed ^ (0xA0761D6478BD642Full * (comp + 1))) ^ (i * 0xE7037ED1A0B428DBull));
}

// One galaxy of n particles: disk 25%, bulge 10% (rounded), halo the rest. Writes
// pos/vel/mass at offset; returns in the galaxy's zero-COM frame (body frame).
void sample_galaxy(const GalaxyModel& gm, size_t n, uint64_t seed, double* pos, double* vel,
                   double* mass) {
  const size_t nd = (size_t)llround(0.25 * (double)n);
  const size_t nb = (size_t)llround(0.10 * (double)n);
  const size_t nh = n - nd - nb;
  const double md = gm.disk_mass * gm.scale / (double)std::max<size_t>(nd, 1);
  const double mb = gm.bulge.m * gm.scale / (double)std::max<size_t>(nb, 1);
  const double mh = gm.halo.m * gm.scale / (double)std::max<size_t>(nh, 1);
  GalaxyModel s = gm;
  s.disk_mass *= gm.scale;
  s.bulge.m *= gm.scale;
  s.halo.m *= gm.scale;
  parallel_for(n, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      V3 p{0, 0, 0}, v{0, 0, 0};
      if (i < nd) {
        // ic/src/disk.rs:333-371 (radius, azimuth, sech^2 height; dispersion about v_c)
        SplitMix64 r(stream_seed(seed, 0, i));
        const double R = s.sample_radius(r.next_f64());
        const double phi = kTau * r.next_f64();
        const double y = r.next_f64();
        const double t = std::min(std::max(2.0 * y - 1.0, -1.0 + 1e-12), 1.0 - 1e-12);
        const double z = s.hz * atanh(t);
        const double cp = cos(phi), sp = sin(phi);
        p = {R * cp, R * sp, z};
        const double vc = s.vcirc(R);
        double gr, gz, gphi, gunused;
        box_muller(r.next_f64(), r.next_f64(), gr, gz);
        box_muller(r.next_f64(), r.next_f64(), gphi, gunused);
        const double sig = gm.sigma_frac * vc;
        const double vr = sig * gr, vphi = vc + sig * gphi, vz = sig * gz;
        v = {vr * cp - vphi * sp, vr * sp + vphi * cp, vz};
        mass[i] = md;
      } else if (i < nd + nb) {
        SplitMix64 r(stream_seed(seed, 1, i));
        s.bulge.sample_one(r, p, v);
        mass[i] = mb;
      } else {
        SplitMix64 r(stream_seed(seed, 2, i));
        s.halo.sample_one(r, p, v);
        mass[i] = mh;
      }
      pos[3 * i] = p.x;
      pos[3 * i + 1] = p.y;
      pos[3 * i + 2] = p.z;
      vel[3 * i] = v.x;
      vel[3 * i + 1] = v.y;
      vel[3 * i + 2] = v.z;
    }
  });
  // mass-weighted recentre (ic/src/disk.rs:380-396)
  double mt = 0, cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
  for (size_t i = 0; i < n; ++i) {
    mt += mass[i];
    for (int c = 0; c < 3; ++c) {
      cp[c] += pos[3 * i + c] * mass[i];
      cv[c] += vel[3 * i + c] * mass[i];
    }
  }
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] -= cp[c] / mt;
      vel[3 * i + c] -= cv[c] / mt;
    }
}

// Rz(arg) * Rx(incl) applied to v (ic/src/orient.rs:63-71)
V3 orient(V3 v, double incl, double arg) {
  const double ci = cos(incl), si = sin(incl), ca = cos(arg), sa = sin(arg);
  const V3 r1{v.x, ci * v.y - si * v.z, si * v.y + ci * v.z};
  return {ca * r1.x - sa * r1.y, sa * r1.x + ca * r1.y, r1.z};
}

}  // namespace

// Two-galaxy collision: n/2 particles per galaxy (disk M=1 R_d=3 sech^2 h=0.3 25%,
// Hernquist bulge M=0.33 a=0.5 10%, Hernquist halo M=5 a=10 65%; disk dispersion
// 10% of v_c), the second galaxy's component masses scaled by mass_ratio, seeded
// random inclinations, parabolic orbit (e = 1) with pericentre rp from separation r0.
extern "C" int glx_ic_collision(size_t n, uint64_t seed, double mass_ratio, double rp, double r0,
                                double* pos, double* vel, double* mass, double* out_info) {
  if (n < 2 || !pos || !vel || !mass || !(mass_ratio > 0) || !(rp > 0) || !(r0 >= rp)) return 1;
  const size_t n1 = n / 2, n2 = n - n1;
  GalaxyModel g1, g2;
  g2.scale = mass_ratio;
  SplitMix64 angles(mix_seed(seed ^ 0x5DEECE66Dull));
  double inc[2], arg[2];
  for (int k = 0; k < 2; ++k) {
    inc[k] = acos(2.0 * angles.next_f64() - 1.0);
    arg[k] = kTau * angles.next_f64();
  }
  sample_galaxy(g1, n1, mix_seed(seed + 1), pos, vel, mass);
  sample_galaxy(g2, n2, mix_seed(seed + 2), pos + 3 * n1, vel + 3 * n1, mass + n1);
  double m1 = 0, m2 = 0;
  for (size_t i = 0; i < n1; ++i) m1 += mass[i];
  for (size_t i = n1; i < n; ++i) m2 += mass[i];
  // ic/src/encounter.rs:45-87, e = 1
  const double mu = (m1 + m2), e = 1.0;
  const double p = rp * (1.0 + e), h = sqrt(mu * p);
  const double cnu = std::min(std::max((p / r0 - 1.0) / e, -1.0), 1.0);
  const double nu = -acos(cnu);
  const double sn = sin(nu), cn = cos(nu);
  const double vr = mu / h * e * sn, vt = mu / h * (1.0 + e * cn);
  const V3 rrel{cn * r0, sn * r0, 0.0};
  const V3 vrel{cn * vr - sn * vt, sn * vr + cn * vt, 0.0};
  const double f1 = -m2 / (m1 + m2), f2 = m1 / (m1 + m2);
  for (size_t i = 0; i < n; ++i) {
    const int k = i < n1 ? 0 : 1;
    const double f = k == 0 ? f1 : f2;
    V3 pp = orient({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, inc[k], arg[k]);
    V3 vv = orient({vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]}, inc[k], arg[k]);
    pos[3 * i] = pp.x + rrel.x * f;
    pos[3 * i + 1] = pp.y + rrel.y * f;
    pos[3 * i + 2] = pp.z + rrel.z * f;
    vel[3 * i] = vv.x + vrel.x * f;
    vel[3 * i + 1] = vv.y + vrel.y * f;
    vel[3 * i + 2] = vv.z + vrel.z * f;
  }
  narrow_f32(pos, 3 * n);
  narrow_f32(vel, 3 * n);
  if (out_info) {
    out_info[0] = inc[0];
    out_info[1] = arg[0];
    out_info[2] = inc[1];
    out_info[3] = arg[1];
  }
  return 0;
}
