/*
 * CPU restatement of the workload generators with one random stream per particle.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device generators (glx_ic_generate_device,
 * glx_pool_admit_ic, glx_sys_generate) and for the host entry point glx_ic_generate. Only
 * tests/ load it. Written as plain serial C: each part (a sphere, or one galaxy) is drawn
 * particle by particle and recentred with a running sum in index order, the direct
 * restatement of the reference's fold (ic/src/plummer.rs:127-135); the library sums in
 * blocks, so the two agree to rounding of the centre, not bit for bit.
 *
 * The samplers restate the reference's ic crate:
 *   SplitMix64, directions      ic/src/plummer.rs:143-174, ic/src/disk.rs:463-479
 *   Plummer AHW                 ic/src/plummer.rs:94-139 (X on [0, M(<rmax)/M))
 *   Hernquist DF and rejection  ic/src/hernquist.rs:78-96,124-176 (cut at 300a)
 *   exponential sech^2 disk     ic/src/disk.rs:270-290,333-371,418-431
 *   Kepler placement            ic/src/encounter.rs:45-87 (e = 1)
 *   orientation Rz(w) Rx(i)     ic/src/orient.rs:63-71
 * and the particle streams are keyed as the library keys them: (seed, component, index)
 * through two rounds of the SplitMix64 finaliser.
 *
 * Pinning (tests/test_ic_streams.py, what the mathematics fixes, after the reference's
 * own sampler tests):
 *   Plummer      enclosed-mass fractions at 5 radii (4-sigma binomial bands), half-mass
 *                radius (3%), virial ratio against W = 3 pi/32 (5%), speed distribution
 *                q^2 (1 - q^2)^3.5 (KS), truncation radius.
 *   Hernquist    DF integrates to the closed-form density (3e-3), enclosed-mass
 *                fractions of the 300a-truncated profile, half-mass radius, virial ratio
 *                against W = 1/6.
 *   galaxy       disk surface density and sech^2 height fractions, v_phi about the
 *                closed-form circular speed with 10% dispersion, bulge and halo profiles,
 *                component masses.
 *   collision    parabolic relative orbit (zero energy, h^2 = 2 mu r_p, approaching),
 *                zero COM and momentum, disk spin along (sin i sin w, -sin i cos w, cos i).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

static const double PI = 3.14159265358979323846;

typedef struct {
  uint64_t s;
} rng_t;

/* ic/src/plummer.rs:153-174 */
static uint64_t rng_u64(rng_t *r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double rng_f64(rng_t *r) { return (double)(rng_u64(r) >> 11) * (1.0 / 9007199254740992.0); }

/* ic/src/disk.rs:474-479: one SplitMix64 step from `seed` */
static uint64_t finalise(uint64_t seed) {
  rng_t r = {seed};
  return rng_u64(&r);
}

static rng_t stream(uint64_t seed, uint64_t comp, uint64_t i) {
  rng_t r;
  r.s = finalise(finalise(seed ^ (0xA0761D6478BD642Full * (comp + 1))) ^ (i * 0xE7037ED1A0B428DBull));
  return r;
}

enum { DISK = 0, BULGE = 1, HALO = 2, PLUMMER = 3 };

/* ic/src/plummer.rs:143-148: cos(theta) uniform, then phi */
static void direction(rng_t *r, double d[3]) {
  double ct = 2.0 * rng_f64(r) - 1.0;
  double s2 = 1.0 - ct * ct;
  double st = sqrt(s2 > 0.0 ? s2 : 0.0);
  double ph = 2.0 * PI * rng_f64(r);
  d[0] = st * cos(ph);
  d[1] = st * sin(ph);
  d[2] = ct;
}

/* Hernquist isotropic DF, ic/src/hernquist.rs:78-96 */
double orc_ic_hernquist_df(double g, double m, double a, double e) {
  if (e <= 0.0) return 0.0;
  double vg2 = g * m / a;
  double q = e / vg2;
  if (q >= 1.0) return 0.0;
  double vg = sqrt(vg2);
  double pref = m / (8.0 * sqrt(2.0) * PI * PI * PI * a * a * a * vg * vg * vg);
  double br = 3.0 * asin(sqrt(q)) + sqrt(q * (1.0 - q)) * (1.0 - 2.0 * q) * (8.0 * q * q - 8.0 * q - 3.0);
  return pref * br / pow(1.0 - q, 2.5);
}

/* One Plummer particle (ic/src/plummer.rs:103-123). */
static void plummer_one(rng_t *r, double g, double m, double a, double x_max, double p[3], double v[3]) {
  double x = x_max * rng_f64(r);
  double rad = a / sqrt(pow(x, -2.0 / 3.0) - 1.0);
  double d[3];
  direction(r, d);
  for (int c = 0; c < 3; ++c) p[c] = d[c] * rad;
  double q = 0.0;
  for (int k = 0; k < 65536; ++k) {
    double qq = rng_f64(r), h = 0.1 * rng_f64(r);
    if (h < qq * qq * pow(1.0 - qq * qq, 3.5)) {
      q = qq;
      break;
    }
  }
  double ve = sqrt(2.0 * g * m) * pow(rad * rad + a * a, -0.25);
  direction(r, d);
  for (int c = 0; c < 3; ++c) v[c] = d[c] * (q * ve);
}

/* One Hernquist particle (ic/src/hernquist.rs:139-176). */
static void hernquist_one(rng_t *r, double g, double m, double a, double p[3], double v[3]) {
  double x_max = (300.0 / 301.0) * (300.0 / 301.0);
  double rx = sqrt(x_max * rng_f64(r));
  double rad = a * rx / (1.0 - rx);
  double d[3];
  direction(r, d);
  for (int c = 0; c < 3; ++c) p[c] = d[c] * rad;
  double psi = g * m / (rad + a), vesc = sqrt(2.0 * psi);
  double top = 0.0;
  for (int k = 1; k < 256; ++k) {
    double vv = vesc * k / 256.0;
    double f = vv * vv * orc_ic_hernquist_df(g, m, a, psi - 0.5 * vv * vv);
    if (f > top) top = f;
  }
  top *= 1.5;
  double sp = 0.0;
  for (int k = 0; k < 65536; ++k) {
    double vv = vesc * rng_f64(r), h = top * rng_f64(r);
    if (h < vv * vv * orc_ic_hernquist_df(g, m, a, psi - 0.5 * vv * vv)) {
      sp = vv;
      break;
    }
  }
  direction(r, d);
  for (int c = 0; c < 3; ++c) v[c] = d[c] * sp;
}

/* Collision galaxy model (disk 1, R_d 3, h_z 0.3, cut 12; bulge 0.33/0.5; halo 5/10; all
 * masses times scale). ic/src/disk.rs:270-288 */
typedef struct {
  double scale;
} galaxy_t;

static double hq_enclosed(double m, double a, double r) { return m * r * r / ((r + a) * (r + a)); }

static double disk_enclosed(const galaxy_t *gx, double r) {
  double x = (r < 12.0 ? r : 12.0) / 3.0;
  double u = 12.0 / 3.0;
  return 1.0 * gx->scale * (1.0 - (1.0 + x) * exp(-x)) / (1.0 - (1.0 + u) * exp(-u));
}

static double vcirc(const galaxy_t *gx, double r) {
  if (r <= 0.0) return 0.0;
  double menc = hq_enclosed(0.33 * gx->scale, 0.5, r) + hq_enclosed(5.0 * gx->scale, 10.0, r) +
                disk_enclosed(gx, r);
  return sqrt(1.0 * menc / r);
}

/* ic/src/disk.rs:463-467 */
static void gauss_pair(double u1, double u2, double *g1, double *g2) {
  double rad = sqrt(-2.0 * log(u1 > 2.2250738585072014e-308 ? u1 : 2.2250738585072014e-308));
  *g1 = rad * cos(2.0 * PI * u2);
  *g2 = rad * sin(2.0 * PI * u2);
}

/* Galaxy of n particles in its body frame, before recentring: disk 25%, bulge 10%
 * (rounded half away from zero), halo the rest. */
void orc_ic_galaxy_body(size_t n, uint64_t seed, double scale, double *pos, double *vel,
                        double *mass) {
  galaxy_t gx = {scale};
  size_t nd = (size_t)llround(0.25 * (double)n), nb = (size_t)llround(0.10 * (double)n);
  for (size_t i = 0; i < n; ++i) {
    double *p = pos + 3 * i, *v = vel + 3 * i;
    if (i < nd) {
      /* ic/src/disk.rs:333-371; radius by 60 bisections of M_d(<R) (ic/src/disk.rs:418-431) */
      rng_t r = stream(seed, DISK, i);
      double target = rng_f64(&r) * (1.0 * scale), lo = 0.0, hi = 12.0;
      for (int k = 0; k < 60; ++k) {
        double mid = 0.5 * (lo + hi);
        if (disk_enclosed(&gx, mid) < target)
          lo = mid;
        else
          hi = mid;
      }
      double R = 0.5 * (lo + hi);
      double phi = 2.0 * PI * rng_f64(&r);
      double t = 2.0 * rng_f64(&r) - 1.0;
      if (t < -1.0 + 1e-12) t = -1.0 + 1e-12;
      if (t > 1.0 - 1e-12) t = 1.0 - 1e-12;
      double cp = cos(phi), sp = sin(phi);
      p[0] = R * cp;
      p[1] = R * sp;
      p[2] = 0.3 * atanh(t);
      double vc = vcirc(&gx, R), gr, gz, gphi, spare;
      double u1 = rng_f64(&r), u2 = rng_f64(&r);
      gauss_pair(u1, u2, &gr, &gz);
      u1 = rng_f64(&r);
      u2 = rng_f64(&r);
      gauss_pair(u1, u2, &gphi, &spare);
      double sig = 0.1 * vc;
      double vr = sig * gr, vphi = vc + sig * gphi;
      v[0] = vr * cp - vphi * sp;
      v[1] = vr * sp + vphi * cp;
      v[2] = sig * gz;
      mass[i] = 1.0 * scale / (double)nd;
    } else if (i < nd + nb) {
      rng_t r = stream(seed, BULGE, i);
      hernquist_one(&r, 1.0, 0.33 * scale, 0.5, p, v);
      mass[i] = 0.33 * scale / (double)nb;
    } else {
      rng_t r = stream(seed, HALO, i);
      hernquist_one(&r, 1.0, 5.0 * scale, 10.0, p, v);
      mass[i] = 5.0 * scale / (double)(n - nd - nb);
    }
  }
}

/* Running mass-weighted centre (index order) of n particles: out = {M, com[3], cov[3]}. */
static void centre(size_t n, const double *pos, const double *vel, const double *mass, double out[7]) {
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < n; ++i) {
    s[0] += mass[i];
    for (int c = 0; c < 3; ++c) {
      s[1 + c] += mass[i] * pos[3 * i + c];
      s[4 + c] += mass[i] * vel[3 * i + c];
    }
  }
  out[0] = s[0];
  for (int c = 1; c < 7; ++c) out[c] = s[c] / s[0];
}

static double f32(double x) { return (double)(float)x; }

/* kind 0: Plummer (rmax <= 0: untruncated), kind 1: Hernquist; recentred, rounded to f32
 * values. */
void orc_ic_sphere(int kind, size_t n, uint64_t seed, double g, double m, double a, double rmax,
                   double *pos, double *vel, double *mass) {
  double x_max = 1.0;
  if (rmax > 0.0 && isfinite(rmax)) x_max = rmax * rmax * rmax / pow(rmax * rmax + a * a, 1.5);
  for (size_t i = 0; i < n; ++i) {
    if (kind == 0) {
      rng_t r = stream(seed, PLUMMER, i);
      plummer_one(&r, g, m, a, x_max, pos + 3 * i, vel + 3 * i);
    } else {
      rng_t r = stream(seed, HALO, i);
      hernquist_one(&r, g, m, a, pos + 3 * i, vel + 3 * i);
    }
    mass[i] = m / (double)n;
  }
  double c[7];
  centre(n, pos, vel, mass, c);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      pos[3 * i + k] = f32(pos[3 * i + k] - c[1 + k]);
      vel[3 * i + k] = f32(vel[3 * i + k] - c[4 + k]);
    }
}

/* ic/src/orient.rs:63-71: Rx(incl) first, then Rz(arg) */
static void rotate(double v[3], double incl, double arg) {
  double y = cos(incl) * v[1] - sin(incl) * v[2];
  double z = sin(incl) * v[1] + cos(incl) * v[2];
  double x = cos(arg) * v[0] - sin(arg) * y;
  double yy = sin(arg) * v[0] + cos(arg) * y;
  v[0] = x;
  v[1] = yy;
  v[2] = z;
}

/* Two galaxies of n/2 and n - n/2 particles on a parabolic orbit with pericentre rp,
 * starting at separation r0, incoming branch; info = {incl1, arg1, incl2, arg2}. */
int orc_ic_collision(size_t n, uint64_t seed, double mass_ratio, double rp, double r0, double *pos,
                     double *vel, double *mass, double *info) {
  if (n < 2 || !(mass_ratio > 0.0) || !(rp > 0.0) || !(r0 >= rp)) return 1;
  size_t cnt[2] = {n / 2, n - n / 2}, off[2] = {0, n / 2};
  rng_t ang = {finalise(seed ^ 0x5DEECE66Dull)};
  double incl[2], arg[2], c[2][7];
  for (int k = 0; k < 2; ++k) {
    incl[k] = acos(2.0 * rng_f64(&ang) - 1.0);
    arg[k] = 2.0 * PI * rng_f64(&ang);
  }
  for (int k = 0; k < 2; ++k) {
    orc_ic_galaxy_body(cnt[k], finalise(seed + 1 + (uint64_t)k), k ? mass_ratio : 1.0,
                       pos + 3 * off[k], vel + 3 * off[k], mass + off[k]);
    centre(cnt[k], pos + 3 * off[k], vel + 3 * off[k], mass + off[k], c[k]);
  }
  /* ic/src/encounter.rs:45-87 with e = 1, G = 1 */
  double mu = c[0][0] + c[1][0], p = 2.0 * rp, h = sqrt(mu * p);
  double cnu = p / r0 - 1.0;
  if (cnu < -1.0) cnu = -1.0;
  if (cnu > 1.0) cnu = 1.0;
  double nu = -acos(cnu), sn = sin(nu), cn = cos(nu);
  double vr = mu / h * sn, vt = mu / h * (1.0 + cn);
  double rrel[3] = {cn * r0, sn * r0, 0.0}, vrel[3] = {cn * vr - sn * vt, sn * vr + cn * vt, 0.0};
  double frac[2] = {-c[1][0] / mu, c[0][0] / mu};
  for (int k = 0; k < 2; ++k)
    for (size_t j = 0; j < cnt[k]; ++j) {
      double *pp = pos + 3 * (off[k] + j), *vv = vel + 3 * (off[k] + j);
      for (int q = 0; q < 3; ++q) {
        pp[q] -= c[k][1 + q];
        vv[q] -= c[k][4 + q];
      }
      rotate(pp, incl[k], arg[k]);
      rotate(vv, incl[k], arg[k]);
      for (int q = 0; q < 3; ++q) {
        pp[q] = f32(pp[q] + rrel[q] * frac[k]);
        vv[q] = f32(vv[q] + vrel[q] * frac[k]);
      }
    }
  if (info) {
    info[0] = incl[0];
    info[1] = arg[0];
    info[2] = incl[1];
    info[3] = arg[1];
  }
  return 0;
}
