/*
 * lbvh_oracle.c — TEST INFRASTRUCTURE ONLY. CPU restatement of the reference's
 * f64 LBVH gravity path, used as the parity checker for the HIP product path and
 * as the timed CPU baseline. Nothing in the product links or calls this file;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function restates a reference function, cited as file:line inside the
 * reference tree (BoykoNeov/galaxy). The f64 operation order of each function
 * follows the cited Rust source (which never contracts to FMA), and this file is
 * compiled with -ffp-contract=off -fno-fast-math so that the bits agree.
 *
 * Pinning (how each function is checked beyond restating it):
 *   expand10/morton3        known answers solvers/src/lbvh.rs:691-711, brute-force
 *                           bit loop over all 1024 lane values, monotone along x.
 *   reference_morton        independent numpy float64 restatement (bit-equal),
 *                           collinear/single/coincident known answers.
 *   reference_sort          numpy argsort(kind="stable") over codes; identity,
 *                           reversed, all-equal known answers.
 *   reference_karras        brute-force top-down split of the augmented 64-bit keys
 *                           (an independent construction of the same radix tree);
 *                           N=1, N=2 known answers; structural invariants.
 *   reference_aggregate     brute-force subtree sums (mass/com to 1e-12) and exact
 *                           subtree min/max (AABB bit-equal).
 *   reference_flatten       invariants (strict increase of next, binary layout,
 *                           AABB containment), N=1/N=2 known answers.
 *   lbvh accelerations      theta->0 equals the direct sum to 1e-9 (worst, RMS
 *                           normalised), momentum flux at theta->0, finite-theta
 *                           error bounds that grow with theta.
 *   direct sum / potential  hand values for two bodies; Plummer closed-form
 *                           W = -3*pi*G*M^2/(32a) on a large sample.
 *   leapfrog KDK            two-body energy conservation, time reversibility.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

#define MORTON_BITS 10u
#define NO_PARENT 0xFFFFFFFFu

/* solvers/src/lbvh.rs:37-44 */
ORC_API uint32_t orc_expand10(uint32_t v) {
    uint32_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

/* solvers/src/lbvh.rs:48-50 */
ORC_API uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) {
    return orc_expand10(x) | (orc_expand10(y) << 1) | (orc_expand10(z) << 2);
}

/* Rust f64::min / f64::max (IEEE minNum/maxNum: a NaN operand yields the other). */
static inline double rmin(double a, double b) { return fmin(a, b); }
static inline double rmax(double a, double b) { return fmax(a, b); }

/* solvers/src/lbvh.rs:58: (v.floor().max(0.0) as u32).min(cells-1).
 * Rust's float->u32 `as` saturates (NaN -> 0, >= 2^32 -> u32::MAX); f64::max(NaN, 0)
 * is already 0, so only the upper saturation needs care. */
static inline uint32_t quant_lane(double v) {
    double f = rmax(floor(v), 0.0);
    if (f >= 4294967295.0) return (1u << MORTON_BITS) - 1u;
    uint32_t u = (uint32_t)f;
    return u < (1u << MORTON_BITS) - 1u ? u : (1u << MORTON_BITS) - 1u;
}

/* solvers/src/lbvh.rs:105-129 (reference_morton) and :56-64 (morton_lanes).
 * pos is interleaved xyz (3n doubles). bounds_out = {bmin.x, bmin.y, bmin.z, size,
 * scale}. lanes (3n) may be NULL. n must be > 0. */
ORC_API int orc_reference_morton(size_t n, const double *pos, double *bounds_out,
                                 uint32_t *lanes, uint32_t *codes) {
    if (n == 0) return -1;
    double lo[3] = {pos[0], pos[1], pos[2]};
    double hi[3] = {pos[0], pos[1], pos[2]};
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = rmin(lo[a], pos[3 * i + a]);
            hi[a] = rmax(hi[a], pos[3 * i + a]);
        }
    }
    double center[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        center[a] = (lo[a] + hi[a]) * 0.5;
        ext[a] = hi[a] - lo[a];
    }
    /* glam DVec3::max_element = x.max(y.max(z)) */
    double maxe = rmax(ext[0], rmax(ext[1], ext[2]));
    double half = rmax(0.5 * maxe, 1e-12) * (1.0 + 1e-9);
    double bmin[3];
    for (int a = 0; a < 3; ++a) bmin[a] = center[a] - half;
    double size = 2.0 * half;
    double scale = (double)(1u << MORTON_BITS) / size;
    if (bounds_out) {
        bounds_out[0] = bmin[0];
        bounds_out[1] = bmin[1];
        bounds_out[2] = bmin[2];
        bounds_out[3] = size;
        bounds_out[4] = scale;
    }
    for (size_t i = 0; i < n; ++i) {
        uint32_t l[3];
        for (int a = 0; a < 3; ++a) l[a] = quant_lane((pos[3 * i + a] - bmin[a]) * scale);
        if (lanes) {
            lanes[3 * i + 0] = l[0];
            lanes[3 * i + 1] = l[1];
            lanes[3 * i + 2] = l[2];
        }
        codes[i] = orc_morton3(l[0], l[1], l[2]);
    }
    return 0;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* solvers/src/lbvh.rs:143-147: order sorted by the unique pair (code, index). */
ORC_API int orc_reference_sort(size_t n, const uint32_t *codes, uint32_t *order) {
    if (n == 0) return 0;
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!k) return -2;
    for (size_t i = 0; i < n; ++i) k[i] = ((uint64_t)codes[i] << 32) | (uint64_t)i;
    qsort(k, n, sizeof(uint64_t), cmp_u64);
    for (size_t i = 0; i < n; ++i) order[i] = (uint32_t)(k[i] & 0xFFFFFFFFu);
    free(k);
    return 0;
}

static inline int64_t clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

/* solvers/src/lbvh.rs:153-165 */
static inline int64_t delta(const uint32_t *codes, int64_t n, int64_t a, int64_t b) {
    if (b < 0 || b >= n) return -1;
    uint32_t ca = codes[a], cb = codes[b];
    if (ca == cb) return 32 + clz32((uint32_t)a ^ (uint32_t)b);
    return clz32(ca ^ cb);
}

/* solvers/src/lbvh.rs:195-240 (karras_node) + :276-306 (reference_karras).
 * left/right have n-1 entries, parent 2n-1 (unified index: leaves [0,n),
 * internal i at n+i, root = n; parent[root] = NO_PARENT). */
ORC_API int orc_reference_karras(size_t n_, const uint32_t *codes, uint32_t *left,
                                 uint32_t *right, uint32_t *parent) {
    if (n_ == 0) return -1;
    int64_t n = (int64_t)n_;
    for (int64_t u = 0; u < 2 * n - 1; ++u) parent[u] = NO_PARENT;
    for (int64_t i = 0; i < n - 1; ++i) {
        int64_t dir = delta(codes, n, i, i + 1) > delta(codes, n, i, i - 1) ? 1 : -1;
        int64_t delta_min = delta(codes, n, i, i - dir);
        int64_t l_max = 2;
        while (delta(codes, n, i, i + l_max * dir) > delta_min) l_max *= 2;
        int64_t l = 0;
        for (int64_t t = l_max / 2; t > 0; t /= 2)
            if (delta(codes, n, i, i + (l + t) * dir) > delta_min) l += t;
        int64_t j = i + l * dir;
        int64_t delta_node = delta(codes, n, i, j);
        int64_t s = 0, t = l;
        for (;;) {
            t = (t + 1) / 2;
            if (delta(codes, n, i, i + (s + t) * dir) > delta_node) s += t;
            if (t <= 1) break;
        }
        int64_t gamma = i + s * dir + (dir < 0 ? dir : 0);
        int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        uint32_t lu = (lo == gamma) ? (uint32_t)gamma : (uint32_t)(n + gamma);
        uint32_t ru = (hi == gamma + 1) ? (uint32_t)(gamma + 1) : (uint32_t)(n + gamma + 1);
        left[i] = lu;
        right[i] = ru;
        parent[lu] = (uint32_t)(n + i);
        parent[ru] = (uint32_t)(n + i);
    }
    return 0;
}

/* ---------------------------------------------------------------------------------
 * 63-bit Morton variant: the reference's deferred resolution refinement
 * (solvers/src/lbvh.rs:29-31 "63-bit (2^21 per axis, two-word sort on the GPU)").
 * Same root-cube convention as reference_morton (:110-120) with cells = 2^21 per axis;
 * the key pair is (code63, index) and delta extends into the index bits past 64.
 * --------------------------------------------------------------------------------- */

/* spread the low 21 bits of v three apart */
ORC_API uint64_t orc_expand21(uint64_t v) {
    uint64_t x = v & 0x1fffffull;
    uint64_t out = 0;
    for (int b = 0; b < 21; ++b) out |= ((x >> b) & 1ull) << (3 * b);
    return out;
}

static inline uint32_t quant_lane21(double v) {
    double f = rmax(floor(v), 0.0);
    const double top = (double)((1u << 21) - 1u);
    return f >= top ? (1u << 21) - 1u : (uint32_t)f;
}

ORC_API int orc_reference_morton63(size_t n, const double *pos, double *bounds_out,
                                   uint32_t *lanes, uint64_t *codes) {
    if (n == 0) return -1;
    double lo[3] = {pos[0], pos[1], pos[2]};
    double hi[3] = {pos[0], pos[1], pos[2]};
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = rmin(lo[a], pos[3 * i + a]);
            hi[a] = rmax(hi[a], pos[3 * i + a]);
        }
    }
    double center[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        center[a] = (lo[a] + hi[a]) * 0.5;
        ext[a] = hi[a] - lo[a];
    }
    double maxe = rmax(ext[0], rmax(ext[1], ext[2]));
    double half = rmax(0.5 * maxe, 1e-12) * (1.0 + 1e-9);
    double bmin[3];
    for (int a = 0; a < 3; ++a) bmin[a] = center[a] - half;
    double size = 2.0 * half;
    double scale = (double)(1u << 21) / size;
    if (bounds_out) {
        bounds_out[0] = bmin[0];
        bounds_out[1] = bmin[1];
        bounds_out[2] = bmin[2];
        bounds_out[3] = size;
        bounds_out[4] = scale;
    }
    for (size_t i = 0; i < n; ++i) {
        uint32_t l[3];
        for (int a = 0; a < 3; ++a) l[a] = quant_lane21((pos[3 * i + a] - bmin[a]) * scale);
        if (lanes) {
            lanes[3 * i + 0] = l[0];
            lanes[3 * i + 1] = l[1];
            lanes[3 * i + 2] = l[2];
        }
        codes[i] = orc_expand21(l[0]) | (orc_expand21(l[1]) << 1) | (orc_expand21(l[2]) << 2);
    }
    return 0;
}

typedef struct {
    uint64_t code;
    uint32_t idx;
} key63_t;

static int cmp_key63(const void *a, const void *b) {
    const key63_t *x = (const key63_t *)a, *y = (const key63_t *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

ORC_API int orc_reference_sort63(size_t n, const uint64_t *codes, uint32_t *order) {
    if (n == 0) return 0;
    key63_t *k = (key63_t *)malloc(n * sizeof(key63_t));
    if (!k) return -2;
    for (size_t i = 0; i < n; ++i) {
        k[i].code = codes[i];
        k[i].idx = (uint32_t)i;
    }
    qsort(k, n, sizeof(key63_t), cmp_key63);
    for (size_t i = 0; i < n; ++i) order[i] = k[i].idx;
    free(k);
    return 0;
}

static inline int64_t delta63(const uint64_t *codes, int64_t n, int64_t a, int64_t b) {
    if (b < 0 || b >= n) return -1;
    uint64_t ca = codes[a], cb = codes[b];
    if (ca == cb) return 64 + clz32((uint32_t)a ^ (uint32_t)b);
    return __builtin_clzll(ca ^ cb);
}

/* solvers/src/lbvh.rs:195-240 over 63-bit codes */
ORC_API int orc_reference_karras63(size_t n_, const uint64_t *codes, uint32_t *left,
                                   uint32_t *right, uint32_t *parent) {
    if (n_ == 0) return -1;
    int64_t n = (int64_t)n_;
    for (int64_t u = 0; u < 2 * n - 1; ++u) parent[u] = NO_PARENT;
    for (int64_t i = 0; i < n - 1; ++i) {
        int64_t dir = delta63(codes, n, i, i + 1) > delta63(codes, n, i, i - 1) ? 1 : -1;
        int64_t delta_min = delta63(codes, n, i, i - dir);
        int64_t l_max = 2;
        while (delta63(codes, n, i, i + l_max * dir) > delta_min) l_max *= 2;
        int64_t l = 0;
        for (int64_t t = l_max / 2; t > 0; t /= 2)
            if (delta63(codes, n, i, i + (l + t) * dir) > delta_min) l += t;
        int64_t j = i + l * dir;
        int64_t delta_node = delta63(codes, n, i, j);
        int64_t s = 0, t = l;
        for (;;) {
            t = (t + 1) / 2;
            if (delta63(codes, n, i, i + (s + t) * dir) > delta_node) s += t;
            if (t <= 1) break;
        }
        int64_t gamma = i + s * dir + (dir < 0 ? dir : 0);
        int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        uint32_t lu = (lo == gamma) ? (uint32_t)gamma : (uint32_t)(n + gamma);
        uint32_t ru = (hi == gamma + 1) ? (uint32_t)(gamma + 1) : (uint32_t)(n + gamma + 1);
        left[i] = lu;
        right[i] = ru;
        parent[lu] = (uint32_t)(n + i);
        parent[ru] = (uint32_t)(n + i);
    }
    return 0;
}

/* One subtree aggregate (solvers/src/lbvh.rs:382-388). */
typedef struct {
    double mass, com[3], mn[3], mx[3];
} agg_t;

/* solvers/src/lbvh.rs:623-636 (fold_agg), fixed (left, right) order. */
static agg_t fold_agg(const agg_t *la, const agg_t *ra) {
    agg_t r;
    r.mass = la->mass + ra->mass;
    for (int a = 0; a < 3; ++a) {
        if (r.mass > 0.0)
            r.com[a] = (la->com[a] * la->mass + ra->com[a] * ra->mass) / r.mass;
        else
            r.com[a] = (la->mn[a] + la->mx[a] + ra->mn[a] + ra->mx[a]) * 0.25;
        r.mn[a] = rmin(la->mn[a], ra->mn[a]);
        r.mx[a] = rmax(la->mx[a], ra->mx[a]);
    }
    return r;
}

/* solvers/src/lbvh.rs:329-379 (reference_aggregate / agg_node). The reference
 * recurses; this walks the same post-order with an explicit stack (no depth limit
 * for monotone chains), producing identical f64 values per node since each node is
 * fold_agg(left, right) of its children's values.
 * Outputs are per unified node (2n-1): aabb_min/aabb_max/com as 3 doubles each. */
ORC_API int orc_reference_aggregate(size_t n_, const uint32_t *left, const uint32_t *right,
                                    const double *sorted_pos, const double *sorted_mass,
                                    double *aabb_min, double *aabb_max, double *com,
                                    double *mass) {
    if (n_ == 0) return -1;
    size_t n = n_, total = 2 * n - 1;
    for (size_t k = 0; k < n; ++k) {
        mass[k] = sorted_mass[k];
        for (int a = 0; a < 3; ++a) {
            double p = sorted_pos[3 * k + a];
            com[3 * k + a] = p;
            aabb_min[3 * k + a] = p;
            aabb_max[3 * k + a] = p;
        }
    }
    if (n == 1) return 0;
    /* post-order over internal nodes: stack of (node, state) */
    uint32_t *stack = (uint32_t *)malloc(total * sizeof(uint32_t));
    uint8_t *state = (uint8_t *)calloc(total, 1);
    if (!stack || !state) {
        free(stack);
        free(state);
        return -2;
    }
    size_t sp = 0;
    stack[sp++] = (uint32_t)n;
    while (sp) {
        uint32_t u = stack[sp - 1];
        size_t i = u - n;
        if (!state[u]) {
            state[u] = 1;
            if (right[i] >= n) stack[sp++] = right[i];
            if (left[i] >= n) stack[sp++] = left[i];
            continue;
        }
        --sp;
        agg_t la, ra;
        uint32_t c[2] = {left[i], right[i]};
        agg_t *dst[2] = {&la, &ra};
        for (int s = 0; s < 2; ++s) {
            dst[s]->mass = mass[c[s]];
            for (int a = 0; a < 3; ++a) {
                dst[s]->com[a] = com[3 * c[s] + a];
                dst[s]->mn[a] = aabb_min[3 * c[s] + a];
                dst[s]->mx[a] = aabb_max[3 * c[s] + a];
            }
        }
        agg_t r = fold_agg(&la, &ra);
        mass[u] = r.mass;
        for (int a = 0; a < 3; ++a) {
            com[3 * u + a] = r.com[a];
            aabb_min[3 * u + a] = r.mn[a];
            aabb_max[3 * u + a] = r.mx[a];
        }
    }
    free(stack);
    free(state);
    return 0;
}

/* solvers/src/lbvh.rs:394-414 (LbvhNode) */
typedef struct {
    double center[3], half[3], com[3], mass, delta;
    uint32_t next, body_start, body_count, pad;
} orc_node_t;

ORC_API size_t orc_node_size(void) { return sizeof(orc_node_t); }

/* solvers/src/lbvh.rs:545-614 (reference_flatten / flatten_unified), with an explicit
 * stack instead of recursion. Children are emitted left-then-right in DFS pre-order;
 * each internal node's `next` is patched once its subtree is laid down. */
ORC_API int orc_reference_flatten(size_t n_, const uint32_t *left, const uint32_t *right,
                                  const double *aabb_min, const double *aabb_max,
                                  const double *com, const double *mass,
                                  const uint32_t *order, orc_node_t *nodes,
                                  uint32_t *leaf_bodies) {
    if (n_ == 0) return -1;
    size_t n = n_, total = 2 * n - 1;
    /* stack entries: unified index with the top bit marking "close" (patch next) */
    uint32_t *stack = (uint32_t *)malloc((total + 1) * sizeof(uint32_t));
    uint32_t *slot = (uint32_t *)malloc(total * sizeof(uint32_t));
    if (!stack || !slot) {
        free(stack);
        free(slot);
        return -2;
    }
    size_t sp = 0, emitted = 0, nbodies = 0;
    stack[sp++] = (n == 1) ? 0u : (uint32_t)n;
    while (sp) {
        uint32_t e = stack[--sp];
        if (e & 0x80000000u) {
            uint32_t u = e & 0x7FFFFFFFu;
            orc_node_t *nd = &nodes[slot[u]];
            double mn[3], mx[3];
            for (int a = 0; a < 3; ++a) {
                mn[a] = aabb_min[3 * u + a];
                mx[a] = aabb_max[3 * u + a];
                nd->center[a] = (mn[a] + mx[a]) * 0.5;
                nd->half[a] = (mx[a] - mn[a]) * 0.5;
                nd->com[a] = com[3 * u + a];
            }
            nd->mass = mass[u];
            double d[3];
            for (int a = 0; a < 3; ++a) d[a] = com[3 * u + a] - nd->center[a];
            nd->delta = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            nd->next = (uint32_t)emitted;
            nd->body_start = 0;
            nd->body_count = 0;
            continue;
        }
        uint32_t u = e;
        size_t me = emitted++;
        orc_node_t *nd = &nodes[me];
        memset(nd, 0, sizeof(*nd));
        if (u < n) {
            for (int a = 0; a < 3; ++a) {
                nd->center[a] = com[3 * u + a];
                nd->com[a] = com[3 * u + a];
            }
            nd->mass = mass[u];
            nd->next = (uint32_t)me + 1;
            nd->body_start = (uint32_t)nbodies;
            nd->body_count = 1;
            leaf_bodies[nbodies++] = order[u];
            continue;
        }
        slot[u] = (uint32_t)me;
        size_t i = u - n;
        stack[sp++] = u | 0x80000000u;
        stack[sp++] = right[i];
        stack[sp++] = left[i];
    }
    free(stack);
    free(slot);
    return 0;
}

/* solvers/src/lbvh.rs:469-524 (LbvhFlat::accel), returning the value before the
 * x G scale. *n_inter counts evaluated terms: accepted internal nodes plus leaf
 * bodies other than the target. */
static void flat_accel(const orc_node_t *nodes, uint32_t n_nodes, const uint32_t *leaf_bodies,
                       size_t target, const double *pos, const double *mass, double theta,
                       double eps2, double *out, uint64_t *n_inter) {
    double xp[3] = {pos[3 * target], pos[3 * target + 1], pos[3 * target + 2]};
    double a[3] = {0.0, 0.0, 0.0};
    uint64_t cnt = 0;
    uint32_t node = 0;
    while (node < n_nodes) {
        const orc_node_t *nd = &nodes[node];
        if (nd->mass <= 0.0) {
            node = nd->next;
            continue;
        }
        if (nd->body_count > 0) {
            uint32_t end = nd->body_start + nd->body_count;
            for (uint32_t k = nd->body_start; k < end; ++k) {
                size_t b = leaf_bodies[k];
                if (b == target) continue;
                double dx[3] = {pos[3 * b] - xp[0], pos[3 * b + 1] - xp[1], pos[3 * b + 2] - xp[2]};
                double r2 = (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) + eps2;
                double w = mass[b] / (r2 * sqrt(r2));
                for (int c = 0; c < 3; ++c) a[c] = a[c] + dx[c] * w;
                ++cnt;
            }
            node = nd->next;
        } else {
            const double *he = nd->half;
            int inside = fabs(xp[0] - nd->center[0]) <= he[0] &&
                         fabs(xp[1] - nd->center[1]) <= he[1] &&
                         fabs(xp[2] - nd->center[2]) <= he[2];
            double dx[3] = {nd->com[0] - xp[0], nd->com[1] - xp[1], nd->com[2] - xp[2]};
            double d2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            double s = 2.0 * rmax(he[0], rmax(he[1], he[2]));
            double d = sqrt(d2);
            if (!inside && theta * (d - nd->delta) >= s) {
                double r2 = d2 + eps2;
                double w = nd->mass / (r2 * sqrt(r2));
                for (int c = 0; c < 3; ++c) a[c] = a[c] + dx[c] * w;
                ++cnt;
                node = nd->next;
            } else {
                node += 1;
            }
        }
    }
    out[0] = a[0];
    out[1] = a[1];
    out[2] = a[2];
    if (n_inter) *n_inter = cnt;
}

/* A built flat tree held by the oracle between the build and the walk. */
typedef struct {
    size_t n;
    orc_node_t *nodes;
    uint32_t *leaf_bodies;
} orc_flat_t;

static void flat_free(orc_flat_t *f) {
    free(f->nodes);
    free(f->leaf_bodies);
    f->nodes = NULL;
    f->leaf_bodies = NULL;
}

/* solvers/src/lbvh.rs:440-463 (LbvhFlat::build): morton -> sort -> gather ->
 * karras -> aggregate -> flatten. */
static int flat_build(size_t n, const double *pos, const double *mass, orc_flat_t *f) {
    f->n = n;
    f->nodes = NULL;
    f->leaf_bodies = NULL;
    size_t total = 2 * n - 1;
    uint32_t *codes = malloc(n * 4), *order = malloc(n * 4), *sc = malloc(n * 4);
    double *sp = malloc(n * 24), *sm = malloc(n * 8);
    uint32_t *left = malloc((n ? n - 1 : 0) * 4 + 4), *right = malloc((n ? n - 1 : 0) * 4 + 4);
    uint32_t *parent = malloc(total * 4);
    double *amin = malloc(total * 24), *amax = malloc(total * 24), *acom = malloc(total * 24);
    double *amass = malloc(total * 8);
    f->nodes = malloc(total * sizeof(orc_node_t));
    f->leaf_bodies = malloc(n * 4);
    int rc = 0;
    if (!codes || !order || !sc || !sp || !sm || !left || !right || !parent || !amin || !amax ||
        !acom || !amass || !f->nodes || !f->leaf_bodies) {
        rc = -2;
        goto done;
    }
    orc_reference_morton(n, pos, NULL, NULL, codes);
    orc_reference_sort(n, codes, order);
    for (size_t k = 0; k < n; ++k) {
        uint32_t i = order[k];
        sc[k] = codes[i];
        sp[3 * k] = pos[3 * i];
        sp[3 * k + 1] = pos[3 * i + 1];
        sp[3 * k + 2] = pos[3 * i + 2];
        sm[k] = mass[i];
    }
    orc_reference_karras(n, sc, left, right, parent);
    orc_reference_aggregate(n, left, right, sp, sm, amin, amax, acom, amass);
    orc_reference_flatten(n, left, right, amin, amax, acom, amass, order, f->nodes,
                          f->leaf_bodies);
done:
    free(codes);
    free(order);
    free(sc);
    free(sp);
    free(sm);
    free(left);
    free(right);
    free(parent);
    free(amin);
    free(amax);
    free(acom);
    free(amass);
    if (rc) flat_free(f);
    return rc;
}

/* solvers/src/lbvh.rs:662-676 (Lbvh::accelerations). Targets are all n particles
 * (targets == NULL, n_targets ignored) or the listed subset; acc_out holds 3 doubles
 * per listed target. nthreads > 1 splits the target loop over OpenMP threads (each
 * target is an independent accumulation, so the values do not depend on it).
 * Returns the total evaluated-term count in *interactions (may be NULL). */
ORC_API int orc_lbvh_accelerations(size_t n, const double *pos, const double *mass, double g,
                                   double softening, double theta, const uint32_t *targets,
                                   size_t n_targets, double *acc_out, uint64_t *interactions,
                                   int nthreads) {
    if (interactions) *interactions = 0;
    if (n == 0) return 0;
    orc_flat_t f;
    int rc = flat_build(n, pos, mass, &f);
    if (rc) return rc;
    double eps2 = softening * softening;
    size_t m = targets ? n_targets : n;
    uint64_t total = 0;
    uint32_t n_nodes = (uint32_t)(2 * n - 1);
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) num_threads(nthreads)
#endif
    for (size_t t = 0; t < m; ++t) {
        size_t i = targets ? targets[t] : t;
        double a[3];
        uint64_t c = 0;
        flat_accel(f.nodes, n_nodes, f.leaf_bodies, i, pos, mass, theta, eps2, a, &c);
        acc_out[3 * t] = a[0] * g;
        acc_out[3 * t + 1] = a[1] * g;
        acc_out[3 * t + 2] = a[2] * g;
        total += c;
    }
    if (interactions) *interactions = total;
    flat_free(&f);
    return 0;
}

/* Export of the built flat tree for structural tests: nodes (2n-1) and leaf_bodies (n). */
ORC_API int orc_lbvh_build_flat(size_t n, const double *pos, const double *mass,
                                orc_node_t *nodes_out, uint32_t *leaf_bodies_out) {
    if (n == 0) return 0;
    orc_flat_t f;
    int rc = flat_build(n, pos, mass, &f);
    if (rc) return rc;
    memcpy(nodes_out, f.nodes, (2 * n - 1) * sizeof(orc_node_t));
    memcpy(leaf_bodies_out, f.leaf_bodies, n * 4);
    flat_free(&f);
    return 0;
}

/* solvers/src/direct_sum.rs:25-45 (DirectSum::accelerations, Newton pairing). */
ORC_API void orc_direct_sum(size_t n, const double *pos, const double *mass, double g,
                            double softening, double *acc) {
    for (size_t i = 0; i < 3 * n; ++i) acc[i] = 0.0;
    double eps2 = softening * softening;
    for (size_t i = 0; i < n; ++i) {
        const double *xi = &pos[3 * i];
        double mi = mass[i];
        for (size_t j = i + 1; j < n; ++j) {
            double dx[3] = {pos[3 * j] - xi[0], pos[3 * j + 1] - xi[1], pos[3 * j + 2] - xi[2]};
            double r2 = (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) + eps2;
            double inv_r3 = 1.0 / (r2 * sqrt(r2));
            double w[3];
            for (int c = 0; c < 3; ++c) w[c] = dx[c] * (g * inv_r3);
            for (int c = 0; c < 3; ++c) {
                acc[3 * i + c] = acc[3 * i + c] + w[c] * mass[j];
                acc[3 * j + c] = acc[3 * j + c] - w[c] * mi;
            }
        }
    }
}

/* Direct sum restricted to a target subset (each target against all n sources,
 * self excluded by index, same per-pair kernel as the direct sum). Used for the
 * sampled FP64 error reference; parallel over targets. */
ORC_API void orc_direct_sum_targets(size_t n, const double *pos, const double *mass, double g,
                                    double softening, const uint32_t *targets, size_t n_targets,
                                    double *acc_out, int nthreads) {
    double eps2 = softening * softening;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (size_t t = 0; t < n_targets; ++t) {
        size_t i = targets[t];
        double a[3] = {0, 0, 0};
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                            pos[3 * j + 2] - pos[3 * i + 2]};
            double r2 = (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) + eps2;
            double inv_r3 = 1.0 / (r2 * sqrt(r2));
            for (int c = 0; c < 3; ++c) a[c] += dx[c] * (g * inv_r3) * mass[j];
        }
        acc_out[3 * t] = a[0];
        acc_out[3 * t + 1] = a[1];
        acc_out[3 * t + 2] = a[2];
    }
}

/* solvers/src/potential.rs:30-41,44-50 (potential_energy_serial / row_potential). */
ORC_API double orc_potential_serial(size_t n, const double *pos, const double *mass, double g,
                                    double softening) {
    double eps2 = softening * softening;
    double u = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double row = 0.0;
        for (size_t j = i + 1; j < n; ++j) {
            double dx[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                            pos[3 * j + 2] - pos[3 * i + 2]};
            double r = sqrt((dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) + eps2);
            row -= g * mass[i] * mass[j] / r;
        }
        u += row;
    }
    return u;
}

/* core/src/diagnostics.rs:8-15,38-68: out = {KE, P.x, P.y, P.z, L.x, L.y, L.z,
 * COM.x, COM.y, COM.z}. */
ORC_API void orc_diagnostics(size_t n, const double *pos, const double *vel,
                             const double *mass, double *out) {
    double ke = 0, p[3] = {0, 0, 0}, l[3] = {0, 0, 0}, w[3] = {0, 0, 0}, mt = 0;
    for (size_t i = 0; i < n; ++i) {
        const double *r = &pos[3 * i], *v = &vel[3 * i];
        double m = mass[i];
        ke += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int c = 0; c < 3; ++c) p[c] += v[c] * m;
        double cr[3] = {r[1] * v[2] - r[2] * v[1], r[2] * v[0] - r[0] * v[2],
                        r[0] * v[1] - r[1] * v[0]};
        for (int c = 0; c < 3; ++c) l[c] += cr[c] * m;
        mt += m;
        for (int c = 0; c < 3; ++c) w[c] += r[c] * m;
    }
    out[0] = ke;
    for (int c = 0; c < 3; ++c) {
        out[1 + c] = p[c];
        out[4 + c] = l[c];
        out[7 + c] = w[c] / mt;
    }
}

/* core/src/integrator.rs:37-78 (LeapfrogKdk::step) driving either the LBVH solver
 * (solver = 0) or the direct sum (solver = 1), for nsteps steps from a fresh
 * (unprimed) integrator. pos/vel are updated in place; acc (3n) receives the cached
 * accelerations left by the last closing kick. */
ORC_API int orc_kdk_run(size_t n, double *pos, double *vel, const double *mass, double g,
                        double softening, double theta, double dt, size_t nsteps, int solver,
                        double *acc, int nthreads) {
    if (n == 0) return 0;
    int rc = 0;
#define ORC_FORCE()                                                                         \
    do {                                                                                    \
        if (solver == 0)                                                                    \
            rc = orc_lbvh_accelerations(n, pos, mass, g, softening, theta, NULL, 0, acc, NULL, \
                                        nthreads);                                          \
        else                                                                                \
            orc_direct_sum(n, pos, mass, g, softening, acc);                                \
    } while (0)
    ORC_FORCE();
    if (rc) return rc;
    double half = 0.5 * dt;
    for (size_t s = 0; s < nsteps; ++s) {
        for (size_t k = 0; k < 3 * n; ++k) vel[k] = vel[k] + acc[k] * half;
        for (size_t k = 0; k < 3 * n; ++k) pos[k] = pos[k] + vel[k] * dt;
        ORC_FORCE();
        if (rc) return rc;
        for (size_t k = 0; k < 3 * n; ++k) vel[k] = vel[k] + acc[k] * half;
    }
#undef ORC_FORCE
    return 0;
}
