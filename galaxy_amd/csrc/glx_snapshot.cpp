// GLXYSNAP snapshot streams (io/src/snapshot.rs): a versioned little-endian header
// followed by the State columns stored one after another. Written at version 4; versions
// 1-3 read back with the defaults of the columns they predate (species collisionless,
// u = 0, formation time -inf). Host code: the stream is assembled in memory and moved
// with one fwrite/fread.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "galaxy_sim.h"
#include "glx_abi_util.h"
#include "glx_snapshot.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "snapshot streams are little-endian");

namespace glx::snap {
namespace {

constexpr char kMagic[8] = {'G', 'L', 'X', 'Y', 'S', 'N', 'A', 'P'};
constexpr uint32_t kVersion = 4;
constexpr uint32_t kMinVersion = 1;
constexpr size_t kMaxStr = 1u << 16;

// Well-formed UTF-8 as Rust's String::from_utf8 accepts it: shortest-form sequences of
// 1-4 bytes, no surrogates (U+D800..U+DFFF), nothing above U+10FFFF.
bool valid_utf8(const unsigned char* s, size_t len) {
  size_t i = 0;
  while (i < len) {
    const unsigned c = s[i];
    if (c < 0x80) {
      ++i;
      continue;
    }
    int extra;
    unsigned cp, lo;
    if (c >= 0xC2 && c <= 0xDF) {
      extra = 1, cp = c & 0x1F, lo = 0x80;
    } else if (c >= 0xE0 && c <= 0xEF) {
      extra = 2, cp = c & 0x0F, lo = 0x800;
    } else if (c >= 0xF0 && c <= 0xF4) {
      extra = 3, cp = c & 0x07, lo = 0x10000;
    } else {
      return false;  // a continuation byte, an overlong lead (C0, C1) or F5..FF
    }
    if (len - i <= (size_t)extra) return false;
    for (int k = 1; k <= extra; ++k) {
      const unsigned t = s[i + k];
      if ((t & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (t & 0x3F);
    }
    if (cp < lo || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) return false;
    i += 1 + extra;
  }
  return true;
}

struct Out {
  std::vector<unsigned char> b;
  template <class T>
  void put(T v) {
    const size_t o = b.size();
    b.resize(o + sizeof(T));
    std::memcpy(b.data() + o, &v, sizeof(T));
  }
  void str(const char* s) {
    const size_t len = s ? std::strlen(s) : 0;
    if (len > kMaxStr) fail(GLX_EINVAL, "header string exceeds 65536 bytes");
    // the reference writes Rust &str values, which are always UTF-8
    if (!valid_utf8(reinterpret_cast<const unsigned char*>(s), len))
      fail(GLX_EINVAL, "header string is not valid UTF-8");
    put<uint32_t>((uint32_t)len);
    b.insert(b.end(), s, s + len);
  }
};

struct In {
  const unsigned char* p;
  size_t left;
  template <class T>
  T get() {
    if (left < sizeof(T)) fail(GLX_EFORMAT, "truncated snapshot stream");
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    left -= sizeof(T);
    return v;
  }
  std::string str() {
    const uint32_t len = get<uint32_t>();
    if (len > kMaxStr) fail(GLX_EFORMAT, "header string length exceeds 65536 bytes");
    if (left < len) fail(GLX_EFORMAT, "truncated snapshot stream");
    // io/src/snapshot.rs read_str decodes with String::from_utf8 and rejects the rest
    if (!valid_utf8(p, len)) fail(GLX_EFORMAT, "header string is not valid UTF-8");
    std::string s(reinterpret_cast<const char*>(p), len);
    p += len;
    left -= len;
    return s;
  }
};

std::vector<unsigned char> slurp(const char* path) {
  if (!path) fail(GLX_EINVAL, "path is NULL");
  std::FILE* f = std::fopen(path, "rb");
  if (!f) fail(GLX_EIO, std::string("cannot open ") + path);
  std::vector<unsigned char> b;
  unsigned char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + got);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) fail(GLX_EIO, std::string("cannot read ") + path);
  return b;
}

}  // namespace

void write(const char* path, const glx_snap_header& h, size_t n, const double* pos,
           const double* vel, const double* mass, const glx_snap_columns* cols) {
  if (!path) fail(GLX_EINVAL, "path is NULL");
  if (n && (!pos || !vel || !mass)) fail(GLX_EINVAL, "NULL array");
  if (n > ((size_t)1 << 40)) fail(GLX_EINVAL, "n too large");
  const glx_snap_columns none{};
  const glx_snap_columns& c = cols ? *cols : none;
  Out o;
  o.b.reserve(80 + n * (24 + 24 + 4 + 8 + 2 + 1 + 8 + 8));
  o.b.insert(o.b.end(), kMagic, kMagic + 8);
  o.put<uint32_t>(kVersion);
  o.put<double>(h.time);
  o.put<uint64_t>(h.step);
  o.put<double>(h.scale_factor);
  o.put<double>(h.softening);
  o.put<uint64_t>((uint64_t)n);  // the count comes from the data
  o.put<uint64_t>(h.rng_seed);
  o.put<uint64_t>(h.config_hash);
  o.str(h.units);
  o.str(h.code_version);
  for (size_t i = 0; i < 3 * n; ++i) o.put<double>(pos[i]);
  for (size_t i = 0; i < 3 * n; ++i) o.put<double>(vel[i]);
  for (size_t i = 0; i < n; ++i) o.put<float>((float)mass[i]);
  for (size_t i = 0; i < n; ++i) o.put<uint64_t>(c.id ? c.id[i] : (uint64_t)i);
  for (size_t i = 0; i < n; ++i) o.put<uint16_t>(c.progenitor ? c.progenitor[i] : 0);
  for (size_t i = 0; i < n; ++i) {
    const uint8_t k = c.kind ? c.kind[i] : 0;
    if (k > 1) fail(GLX_EINVAL, "species must be 0 (collisionless) or 1 (gas)");
    o.put<uint8_t>(k);
  }
  for (size_t i = 0; i < n; ++i) o.put<double>(c.u ? c.u[i] : 0.0);
  for (size_t i = 0; i < n; ++i)
    o.put<double>(c.formation_time ? c.formation_time[i] : -HUGE_VAL);
  std::FILE* f = std::fopen(path, "wb");
  if (!f) fail(GLX_EIO, std::string("cannot create ") + path);
  const bool ok = std::fwrite(o.b.data(), 1, o.b.size(), f) == o.b.size();
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) fail(GLX_EIO, std::string("cannot write ") + path);
}

Parsed read(const char* path, bool header_only) {
  const std::vector<unsigned char> b = slurp(path);
  In in{b.data(), b.size()};
  if (in.left < 8 || std::memcmp(in.p, kMagic, 8) != 0)
    fail(GLX_EFORMAT, "not a galaxy snapshot (bad magic)");
  in.p += 8;
  in.left -= 8;
  Parsed r;
  r.version = in.get<uint32_t>();
  if (r.version < kMinVersion || r.version > kVersion)
    fail(GLX_EFORMAT, "unsupported snapshot format version " + std::to_string(r.version));
  r.time = in.get<double>();
  r.step = in.get<uint64_t>();
  r.scale_factor = in.get<double>();
  r.softening = in.get<double>();
  r.n = in.get<uint64_t>();
  r.rng_seed = in.get<uint64_t>();
  r.config_hash = in.get<uint64_t>();
  r.units = in.str();
  r.code_version = in.str();
  if (header_only) return r;
  const uint64_t n = r.n;
  // per particle: pos 24 + vel 24 + mass 4 + id 8 + progenitor 2 (+ kind 1 + u 8 + ft 8)
  const uint64_t per = 62 + (r.version >= 2 ? 1 : 0) + (r.version >= 3 ? 8 : 0) +
                       (r.version >= 4 ? 8 : 0);
  if (n > in.left / per) fail(GLX_EFORMAT, "truncated snapshot stream");
  r.pos.resize(3 * n);
  r.vel.resize(3 * n);
  r.mass.resize(n);
  r.id.resize(n);
  r.progenitor.resize(n);
  r.kind.assign(n, 0);
  r.u.assign(n, 0.0);
  r.formation_time.assign(n, -HUGE_VAL);
  for (auto& v : r.pos) v = in.get<double>();
  for (auto& v : r.vel) v = in.get<double>();
  for (auto& v : r.mass) v = (double)in.get<float>();
  for (auto& v : r.id) v = in.get<uint64_t>();
  for (auto& v : r.progenitor) v = in.get<uint16_t>();
  if (r.version >= 2)
    for (auto& v : r.kind) {
      v = in.get<uint8_t>();
      if (v > 1) fail(GLX_EFORMAT, "unknown species byte");
    }
  if (r.version >= 3)
    for (auto& v : r.u) v = in.get<double>();
  if (r.version >= 4)
    for (auto& v : r.formation_time) v = in.get<double>();
  return r;
}

}  // namespace glx::snap

using glx::abi::guard;
using glx::snap::Parsed;

namespace {
void copy_str(const std::string& s, char* buf, size_t len, const char** out) {
  if (!buf) {
    *out = nullptr;
    return;
  }
  if (s.size() + 1 > len) glx::fail(GLX_EINVAL, "string buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  *out = buf;
}
}  // namespace

extern "C" int glx_snapshot_write(const char* path, const glx_snap_header* header, size_t n,
                                  const double* pos, const double* vel, const double* mass,
                                  const glx_snap_columns* columns) {
  return guard([&] {
    if (!header) glx::fail(GLX_EINVAL, "header is NULL");
    glx::snap::write(path, *header, n, pos, vel, mass, columns);
  });
}

extern "C" int glx_snapshot_read_header(const char* path, glx_snap_header* header, char* units,
                                        size_t units_len, char* code_version,
                                        size_t code_version_len, uint32_t* version) {
  return guard([&] {
    if (!header) glx::fail(GLX_EINVAL, "header is NULL");
    const Parsed r = glx::snap::read(path, true);
    glx_snap_header h{};
    h.time = r.time;
    h.step = r.step;
    h.scale_factor = r.scale_factor;
    h.softening = r.softening;
    h.n_particles = r.n;
    h.rng_seed = r.rng_seed;
    h.config_hash = r.config_hash;
    copy_str(r.units, units, units_len, &h.units);
    copy_str(r.code_version, code_version, code_version_len, &h.code_version);
    if (version) *version = r.version;
    *header = h;
  });
}

extern "C" int glx_snapshot_read(const char* path, size_t n, double* pos, double* vel,
                                 double* mass, uint64_t* id, uint16_t* progenitor, uint8_t* kind,
                                 double* u, double* formation_time) {
  return guard([&] {
    const Parsed r = glx::snap::read(path, false);
    if (r.n != n) glx::fail(GLX_EINVAL, "n does not match the snapshot header");
    auto put = [](auto* dst, const auto& src) {
      if (dst) std::copy(src.begin(), src.end(), dst);
    };
    put(pos, r.pos);
    put(vel, r.vel);
    put(mass, r.mass);
    put(id, r.id);
    put(progenitor, r.progenitor);
    put(kind, r.kind);
    put(u, r.u);
    put(formation_time, r.formation_time);
  });
}
