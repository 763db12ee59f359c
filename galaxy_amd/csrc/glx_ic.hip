// Initial-condition generators for the benchmark workloads: host entry points and the
// device kernels that draw a system directly into HBM.
//
// The samplers (glx_ic_sampler.h) follow the reference's ic crate. glx_ic_plummer keeps
// the reference's sequential SplitMix64 draw order (ic/src/plummer.rs:94-139); every
// other generator draws each particle from its own stream, so the device computes one
// particle per thread and the host computes the same values.
//
// Device generation of a plan, per part (a galaxy, or the whole sphere):
//   k_ic_moments  one block per kChunk particles: draw each particle, sum its seven
//                 moments (m, m x, m v) over the block in a fixed tree order
//   k_ic_frames   one wave: sum each part's block partials in block order, then the
//                 recentring shift, orientation and Kepler offset of each part
//   k_ic_write    one thread per particle: draw it again, apply its part's frame, round
//                 to f32 values, store
// The host path runs the same three steps in the same orders, so both sides add the
// same numbers in the same sequence.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "galaxy_ic.h"
#include "glx_abi_util.h"
#include "glx_ic_device.h"
#include "glx_ic_sampler.h"

namespace glx {
namespace ic {
namespace {

template <class F>
void parallel_for(size_t n, F&& f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  nt = std::min<unsigned>(nt, 16);
  if (n < 16 || nt == 1) {
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

Galaxy collision_galaxy(uint64_t n, uint64_t seed, double scale) {
  Galaxy g{};
  g.g = 1.0;
  g.scale = scale;
  g.disk_mass = 1.0 * scale;
  g.rd = 3.0;
  g.hz = 0.3;
  g.rmax = 12.0;  // 4 R_d
  g.sigma_frac = 0.1;
  g.bulge = Hernquist{1.0, 0.33 * scale, 0.5};
  g.halo = Hernquist{1.0, 5.0 * scale, 10.0};
  g.n = n;
  g.nd = (uint64_t)llround(0.25 * (double)n);
  g.nb = (uint64_t)llround(0.10 * (double)n);
  g.seed = seed;
  return g;
}

void set_part(Plan& pl, int q, uint64_t off, uint64_t cnt, uint64_t blk0) {
  pl.off[q] = off;
  pl.cnt[q] = cnt;
  pl.blk0[q] = blk0;
  pl.blks[q] = (cnt + kChunk - 1) / kChunk;
}

uint64_t total_blocks(const Plan& pl) {
  return pl.parts == 2 ? pl.blks[0] + pl.blks[1] : pl.blks[0];
}

// Moments of block b of part q, reduced over kChunk slots by halving (slot t += t + s).
void block_moments(const Plan& pl, int q, uint64_t b, double out[kMoments]) {
  double a[kChunk][kMoments];
  for (int t = 0; t < kChunk; ++t) {
    const uint64_t k = b * kChunk + t;
    if (k < pl.cnt[q]) {
      V3 p, v;
      double m;
      plan_sample(pl, q, k, p, v, m);
      moments_of(p, v, m, a[t]);
    } else {
      for (int c = 0; c < kMoments; ++c) a[t][c] = 0.0;
    }
  }
  for (int s = kChunk / 2; s > 0; s >>= 1)
    for (int t = 0; t < s; ++t)
      for (int c = 0; c < kMoments; ++c) a[t][c] += a[t + s][c];
  for (int c = 0; c < kMoments; ++c) out[c] = a[0][c];
}

// The host path: partials, frames, final values, in the device's orders.
void generate_host(const Plan& pl, double* pos, double* vel, double* mass) {
  const uint64_t nb = total_blocks(pl);
  std::vector<double> partial(nb * kMoments);
  for (int q = 0; q < pl.parts; ++q)
    parallel_for(pl.blks[q], [&](size_t b, size_t e) {
      for (size_t k = b; k < e; ++k) block_moments(pl, q, k, &partial[(pl.blk0[q] + k) * kMoments]);
    });
  double sums[2][kMoments] = {};
  for (int q = 0; q < pl.parts; ++q)
    for (uint64_t b = 0; b < pl.blks[q]; ++b)
      for (int c = 0; c < kMoments; ++c) sums[q][c] += partial[(pl.blk0[q] + b) * kMoments + c];
  Frame fr[2];
  make_frames(pl, sums, fr);
  for (int q = 0; q < pl.parts; ++q)
    parallel_for(pl.cnt[q], [&](size_t b, size_t e) {
      for (size_t k = b; k < e; ++k) {
        V3 p, v;
        double m;
        plan_sample(pl, q, k, p, v, m);
        apply_frame(fr[q], p, v);
        const size_t i = pl.off[q] + k;
        pos[3 * i] = p.x;
        pos[3 * i + 1] = p.y;
        pos[3 * i + 2] = p.z;
        vel[3 * i] = v.x;
        vel[3 * i + 1] = v.y;
        vel[3 * i + 2] = v.z;
        mass[i] = m;
      }
    });
}

// ---------------------------------------------------------------------------------
// Device kernels. Grids: k_ic_moments total_blocks(pl) x kChunk; k_ic_frames 1 x 64;
// k_ic_write ceil(n / kChunk) x kChunk. Every index is checked against its part's count.

__global__ void __launch_bounds__(kChunk) k_ic_moments(Plan pl, double* partial) {
  __shared__ double a[kMoments][kChunk];
  const uint64_t blk = blockIdx.x;
  const int q = (pl.parts == 2 && blk >= pl.blk0[1]) ? 1 : 0;
  const uint64_t k = (blk - pl.blk0[q]) * kChunk + threadIdx.x;
  double mo[kMoments] = {0, 0, 0, 0, 0, 0, 0};
  if (k < pl.cnt[q]) {
    V3 p, v;
    double m;
    plan_sample(pl, q, k, p, v, m);
    moments_of(p, v, m, mo);
  }
  for (int c = 0; c < kMoments; ++c) a[c][threadIdx.x] = mo[c];
  __syncthreads();
  for (int s = kChunk / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < kMoments; ++c) a[c][threadIdx.x] += a[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < kMoments) partial[blk * kMoments + threadIdx.x] = a[threadIdx.x][0];
}

__global__ void __launch_bounds__(64) k_ic_frames(Plan pl, const double* partial, Frame* out) {
  __shared__ double sums[2][kMoments];
  const int t = threadIdx.x;
  if (t < 2 * kMoments) {
    const int q = t / kMoments, c = t % kMoments;
    double s = 0.0;
    if (q < pl.parts)
      for (uint64_t b = 0; b < pl.blks[q]; ++b) s += partial[(pl.blk0[q] + b) * kMoments + c];
    sums[q][c] = s;
  }
  __syncthreads();
  if (t == 0) {
    Frame fr[2];
    make_frames(pl, sums, fr);
    for (int q = 0; q < pl.parts; ++q) out[q] = fr[q];
  }
}

template <class VelT>
__global__ void __launch_bounds__(kChunk) k_ic_write(Plan pl, const Frame* frames, double* pos,
                                                     VelT* vel, double* mass) {
  const uint64_t i = (uint64_t)blockIdx.x * kChunk + threadIdx.x;
  if (i >= pl.n) return;
  const int q = (pl.parts == 2 && i >= pl.off[1]) ? 1 : 0;
  const uint64_t k = i - pl.off[q];
  if (k >= pl.cnt[q]) return;
  V3 p, v;
  double m;
  plan_sample(pl, q, k, p, v, m);
  apply_frame(frames[q], p, v);
  pos[3 * i] = p.x;
  pos[3 * i + 1] = p.y;
  pos[3 * i + 2] = p.z;
  vel[3 * i] = (VelT)v.x;
  vel[3 * i + 1] = (VelT)v.y;
  vel[3 * i + 2] = (VelT)v.z;
  mass[i] = m;
}

template <class VelT>
void generate_dev(const Plan& pl, double* pos, VelT* vel, double* mass, hipStream_t st) {
  const uint64_t nb = total_blocks(pl);
  abi::DevArray<double> partial(nb * kMoments);
  abi::DevArray<Frame> frames(2);
  k_ic_moments<<<dim3((unsigned)nb), dim3(kChunk), 0, st>>>(pl, partial.p);
  check(hipGetLastError(), "k_ic_moments");
  k_ic_frames<<<1, 64, 0, st>>>(pl, partial.p, frames.p);
  check(hipGetLastError(), "k_ic_frames");
  const uint64_t wb = (pl.n + kChunk - 1) / kChunk;
  k_ic_write<VelT><<<dim3((unsigned)wb), dim3(kChunk), 0, st>>>(pl, frames.p, pos, vel, mass);
  check(hipGetLastError(), "k_ic_write");
  abi::sync(st);  // partial and frames are freed on return
}

}  // namespace

Plan make_plan(const glx_ic_spec& s) {
  if (s.n > abi::kMaxN) fail(GLX_EINVAL, "n exceeds 2^29 particles per system");
  Plan pl{};
  pl.kind = s.kind;
  pl.n = s.n;
  pl.seed = s.seed;
  if (s.kind == GLX_IC_PLUMMER || s.kind == GLX_IC_HERNQUIST) {
    if (s.n == 0) fail(GLX_EINVAL, "a system needs at least one particle");
    if (!(s.g > 0) || !(s.m > 0) || !(s.a > 0) || !isfinite(s.g) || !isfinite(s.m) ||
        !isfinite(s.a))
      fail(GLX_EINVAL, "g, m and a must be positive and finite");
    pl.parts = 1;
    set_part(pl, 0, 0, s.n, 0);
    if (s.kind == GLX_IC_PLUMMER) {
      const double x_max = (s.rmax > 0 && isfinite(s.rmax))
                               ? s.rmax * s.rmax * s.rmax / pow(s.rmax * s.rmax + s.a * s.a, 1.5)
                               : 1.0;
      pl.plummer = Plummer{s.g, s.m, s.a, x_max};
    } else {
      pl.hernquist = Hernquist{s.g, s.m, s.a};
    }
    return pl;
  }
  if (s.kind != GLX_IC_COLLISION) fail(GLX_EINVAL, "unknown generator kind");
  if (s.n < 2) fail(GLX_EINVAL, "a collision needs at least two particles");
  if (!(s.mass_ratio > 0) || !isfinite(s.mass_ratio) || !(s.rp > 0) || !isfinite(s.r0) ||
      !(s.r0 >= s.rp))
    fail(GLX_EINVAL, "need mass_ratio > 0 and 0 < rp <= r0");
  const uint64_t n1 = s.n / 2, n2 = s.n - n1;
  pl.parts = 2;
  set_part(pl, 0, 0, n1, 0);
  set_part(pl, 1, n1, n2, pl.blks[0]);
  pl.galaxy[0] = collision_galaxy(n1, mix_seed(s.seed + 1), 1.0);
  pl.galaxy[1] = collision_galaxy(n2, mix_seed(s.seed + 2), s.mass_ratio);
  SplitMix64 angles(mix_seed(s.seed ^ 0x5DEECE66Dull));
  for (int k = 0; k < 2; ++k) {
    pl.incl[k] = acos(2.0 * angles.next_f64() - 1.0);
    pl.arg[k] = kTau * angles.next_f64();
  }
  pl.rp = s.rp;
  pl.r0 = s.r0;
  return pl;
}

void generate_resident(const glx_ic_spec& s, double* pos, float* vel, double* mass,
                       hipStream_t st) {
  const Plan pl = make_plan(s);
  generate_dev(pl, pos, vel, mass, st);
}

void validate_spec(const glx_ic_spec& s) { (void)make_plan(s); }

double spec_g(const glx_ic_spec& s) { return s.kind == GLX_IC_COLLISION ? 1.0 : s.g; }

}  // namespace ic
}  // namespace glx

using glx::ic::V3;

// ic/src/plummer.rs:94-139 with truncation at rmax (X uniform on [0, M(<rmax)/M)), in
// the reference's sequential draw order.
extern "C" int glx_ic_plummer(size_t n, uint64_t seed, double g, double m, double a, double rmax,
                              double* pos, double* vel, double* mass) {
  if (n == 0) return 0;
  if (!pos || !vel || !mass || !(g > 0) || !(m > 0) || !(a > 0)) return 1;
  glx::ic::SplitMix64 rng(seed);
  const double x_max = (rmax > 0 && isfinite(rmax))
                           ? rmax * rmax * rmax / pow(rmax * rmax + a * a, 1.5)
                           : 1.0;
  const glx::ic::Plummer pl{g, m, a, x_max};
  for (size_t i = 0; i < n; ++i) {
    V3 p, v;
    pl.sample_one(rng, p, v);
    pos[3 * i] = p.x;
    pos[3 * i + 1] = p.y;
    pos[3 * i + 2] = p.z;
    vel[3 * i] = v.x;
    vel[3 * i + 1] = v.y;
    vel[3 * i + 2] = v.z;
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
  const double me = m / (double)n;
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] = (double)(float)(pos[3 * i + c] - mp[c]);
      vel[3 * i + c] = (double)(float)(vel[3 * i + c] - mv[c]);
    }
    mass[i] = me;
  }
  return 0;
}

extern "C" int glx_ic_generate(const glx_ic_spec* spec, double* pos, double* vel, double* mass,
                               double* out_info) {
  using namespace glx;
  return abi::guard([&] {
    if (!spec) fail(GLX_EINVAL, "spec is NULL");
    const ic::Plan pl = ic::make_plan(*spec);
    if (!pos || !vel || !mass) fail(GLX_EINVAL, "NULL array");
    ic::generate_host(pl, pos, vel, mass);
    if (out_info) {
      out_info[0] = pl.incl[0];
      out_info[1] = pl.arg[0];
      out_info[2] = pl.incl[1];
      out_info[3] = pl.arg[1];
    }
  });
}

extern "C" int glx_ic_generate_device(const glx_ic_spec* spec, double* pos, double* vel,
                                      double* mass) {
  using namespace glx;
  return abi::guard([&] {
    if (!spec) fail(GLX_EINVAL, "spec is NULL");
    const ic::Plan pl = ic::make_plan(*spec);
    if (!pos || !vel || !mass) fail(GLX_EINVAL, "NULL array");
    ensure_device();
    abi::Stream st;
    ic::generate_dev(pl, pos, vel, mass, st.st);
  });
}

// Two-galaxy collision (see glx_ic_generate with GLX_IC_COLLISION).
extern "C" int glx_ic_collision(size_t n, uint64_t seed, double mass_ratio, double rp, double r0,
                                double* pos, double* vel, double* mass, double* out_info) {
  glx_ic_spec s{};
  s.kind = GLX_IC_COLLISION;
  s.n = n;
  s.seed = seed;
  s.mass_ratio = mass_ratio;
  s.rp = rp;
  s.r0 = r0;
  return glx_ic_generate(&s, pos, vel, mass, out_info) == GLX_OK ? 0 : 1;
}
