/*
 * bh_oracle.c -- CPU restatement of the reference Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see bh_oracle.h).
 * Built with -ffp-contract=off so that every float operation below is one
 * IEEE-754 binary32 operation, evaluated in the order the WGSL writes it.
 * WGSL's inverseSqrt() is restated as 1/sqrt() with both correctly rounded.
 */
#include "bh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

/* WGSL clamp(e, lo, hi) = min(max(e, lo), hi); NaN maps to lo. */
static inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

uint32_t orc_interleave16(uint32_t x, uint32_t y) {
    /* lbvh.wgsl:59-65: bit b of x goes to bit 2b, bit b of y to bit 2b+1. */
    uint32_t code = 0;
    for (uint32_t bit = 0; bit < 16; ++bit) {
        uint32_t mask = 1u << bit;
        uint32_t bx = ((x & mask) >> bit) << (2u * bit);
        uint32_t by = ((y & mask) >> bit) << (2u * bit + 1u);
        code |= bx | by;
    }
    return code;
}

void orc_morton_ref(int64_t n, const float *pos_xy, uint32_t *keys, uint32_t *idx) {
    /* lbvh.wgsl:42-69 */
    const float scale = 2000.0f;
    const float half = 0.5f * scale;
    for (int64_t i = 0; i < n; ++i) {
        float ux = (pos_xy[2 * i] + half) / scale;
        float uy = (pos_xy[2 * i + 1] + half) / scale;
        ux = clampf(ux, 0.0f, 1.0f);
        uy = clampf(uy, 0.0f, 1.0f);
        uint32_t x = (uint32_t)(ux * 65535.0f + 0.5f);
        uint32_t y = (uint32_t)(uy * 65535.0f + 0.5f);
        if (x > 65535u) x = 65535u;
        if (y > 65535u) y = 65535u;
        keys[i] = orc_interleave16(x, y);
        idx[i] = (uint32_t)i;
    }
}

void orc_bbox(int64_t n, const float *pos_xy, float *bbox4) {
    float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        mnx = fminf(mnx, pos_xy[2 * i]);
        mny = fminf(mny, pos_xy[2 * i + 1]);
        mxx = fmaxf(mxx, pos_xy[2 * i]);
        mxy = fmaxf(mxy, pos_xy[2 * i + 1]);
    }
    bbox4[0] = mnx;
    bbox4[1] = mny;
    bbox4[2] = mxx;
    bbox4[3] = mxy;
}

/* Square quantisation box of the bbox30 key mode: origin at the bbox
 * minimum, side = the larger extent (1 when the box is empty, a point, or
 * not finite-positive).  Shared by the plain and ensemble restatements. */
static inline float bbox30_side(const float *bbox4) {
    float side = fmaxf(bbox4[2] - bbox4[0], bbox4[3] - bbox4[1]);
    if (!(side > 0.0f)) side = 1.0f;
    return side;
}

static inline uint32_t bbox30_code(float px, float py, const float *bbox4, float side) {
    float ux = clampf((px - bbox4[0]) / side, 0.0f, 1.0f);
    float uy = clampf((py - bbox4[1]) / side, 0.0f, 1.0f);
    uint32_t x = (uint32_t)(ux * 32767.0f + 0.5f);
    uint32_t y = (uint32_t)(uy * 32767.0f + 0.5f);
    if (x > 32767u) x = 32767u;
    if (y > 32767u) y = 32767u;
    return orc_interleave16(x, y);
}

void orc_morton_bbox30(int64_t n, const float *pos_xy, const float *bbox4, uint32_t *keys,
                       uint32_t *idx) {
    float side = bbox30_side(bbox4);
    for (int64_t i = 0; i < n; ++i) {
        keys[i] = bbox30_code(pos_xy[2 * i], pos_xy[2 * i + 1], bbox4, side);
        idx[i] = (uint32_t)i;
    }
}

/* ------------------------------------------------------------------ sort */

void orc_sort_pairs_u32(int64_t n, uint32_t *keys, uint32_t *vals, int bits) {
    if (n <= 1) return;
    uint32_t *k2 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t *v2 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int shift = 0; shift < bits; shift += 8) {
        int64_t count[257];
        memset(count, 0, sizeof(count));
        for (int64_t i = 0; i < n; ++i) count[((keys[i] >> shift) & 0xFFu) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (int64_t i = 0; i < n; ++i) {
            int64_t dst = count[(keys[i] >> shift) & 0xFFu]++;
            k2[dst] = keys[i];
            v2[dst] = vals[i];
        }
        memcpy(keys, k2, sizeof(uint32_t) * (size_t)n);
        memcpy(vals, v2, sizeof(uint32_t) * (size_t)n);
    }
    free(k2);
    free(v2);
}

void orc_sort_pairs_u64(int64_t n, uint64_t *keys, uint32_t *vals, int bits) {
    if (n <= 1) return;
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t *v2 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int shift = 0; shift < bits; shift += 8) {
        int64_t count[257];
        memset(count, 0, sizeof(count));
        for (int64_t i = 0; i < n; ++i) count[((keys[i] >> shift) & 0xFFu) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (int64_t i = 0; i < n; ++i) {
            int64_t dst = count[(keys[i] >> shift) & 0xFFu]++;
            k2[dst] = keys[i];
            v2[dst] = vals[i];
        }
        memcpy(keys, k2, sizeof(uint64_t) * (size_t)n);
        memcpy(vals, v2, sizeof(uint32_t) * (size_t)n);
    }
    free(k2);
    free(v2);
}

/* ----------------------------------------------------------------- build */

/* lbvh.wgsl:72-83 */
static inline int delta_u32(int64_t i, int64_t j, int64_t n, const uint32_t *k) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

/* Ensemble delta: equal keys share a sim; the tie-break uses sim-local
 * sorted positions so each sim's sub-tree equals its plain tree. */
static inline int delta_u64(int64_t i, int64_t j, int64_t n, const uint64_t *k,
                            const int64_t *sim_offset) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) {
        int64_t off = sim_offset[a >> 30];
        return 64 + clz32((uint32_t)(i - off) ^ (uint32_t)(j - off));
    }
    return clz64(a ^ b);
}

#define KARRAS_BODY(DELTA)                                                        \
    /* lbvh.wgsl:101-161, Karras 2012 Fig. 4 */                                   \
    int d = (DELTA(i, i + 1) > DELTA(i, i - 1)) ? 1 : -1;                         \
    int delta_min = DELTA(i, i - d);                                              \
    int64_t l_max = 2;                                                            \
    while (DELTA(i, i + d * l_max) > delta_min) l_max *= 2;                       \
    int64_t l = 0;                                                                \
    for (int64_t t = l_max / 2; t >= 1; t /= 2)                                   \
        if (DELTA(i, i + d * (l + t)) > delta_min) l += t;                        \
    int64_t j = i + d * l;                                                        \
    int delta_node = DELTA(i, j);                                                 \
    int64_t s = 0;                                                                \
    int64_t t = (l + 1) / 2;                                                      \
    for (;;) {                                                                    \
        if (DELTA(i, i + d * (s + t)) > delta_node) s += t;                       \
        if (t == 1) break;                                                        \
        t = (t + 1) / 2;                                                          \
    }                                                                             \
    int64_t gamma = i + d * s + (d < 0 ? d : 0);                                  \
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;                               \
    uint32_t lc = (lo == gamma) ? (uint32_t)(n - 1 + gamma) : (uint32_t)gamma;    \
    uint32_t rc = (hi == gamma + 1) ? (uint32_t)(n - 1 + gamma + 1) : (uint32_t)(gamma + 1); \
    left[i] = lc;                                                                 \
    right[i] = rc;                                                                \
    parent[lc] = (uint32_t)i;                                                     \
    parent[rc] = (uint32_t)i;

void orc_build_u32(int64_t n, const uint32_t *keys, uint32_t *left, uint32_t *right,
                   uint32_t *parent) {
    if (n <= 0) return;
    parent[0] = 0xFFFFFFFFu;
#define D32(a, b) delta_u32((a), (b), n, keys)
    for (int64_t i = 0; i + 1 < n; ++i) {
        KARRAS_BODY(D32)
    }
#undef D32
    parent[0] = 0xFFFFFFFFu;
}

void orc_build_u64(int64_t n, const uint64_t *keys, const int64_t *sim_offset, int32_t num_sims,
                   uint32_t *left, uint32_t *right, uint32_t *parent, uint32_t *sim_root) {
    for (int32_t s = 0; s < num_sims; ++s) sim_root[s] = 0xFFFFFFFFu;
    if (n <= 0) return;
    parent[0] = 0xFFFFFFFFu;
#define D64(a, b) delta_u64((a), (b), n, keys, sim_offset)
    for (int64_t i = 0; i + 1 < n; ++i) {
        KARRAS_BODY(D64)
        /* A node whose range spans two sims is above the sim level; each of
         * its children whose range lies in one sim is that sim's root. */
        uint64_t s_lo = keys[lo] >> 30, s_hi = keys[hi] >> 30;
        if (s_lo != s_hi) {
            if ((keys[gamma] >> 30) == s_lo) sim_root[s_lo] = lc;
            if ((keys[gamma + 1] >> 30) == s_hi) sim_root[s_hi] = rc;
        }
    }
#undef D64
    parent[0] = 0xFFFFFFFFu;
    if ((keys[0] >> 30) == (keys[n - 1] >> 30)) sim_root[keys[0] >> 30] = 0;
}

/* ------------------------------------------------------------------ fill */

static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* lbvh.wgsl:205-228: merge the two children of p. */
static void merge_node(uint32_t p, const uint32_t *left, const uint32_t *right, float *cm,
                       float *bb) {
    uint32_t l = left[p], r = right[p];
    const float *L = cm + 4 * (size_t)l, *R = cm + 4 * (size_t)r;
    const float *LB = bb + 4 * (size_t)l, *RB = bb + 4 * (size_t)r;
    float *P = cm + 4 * (size_t)p, *PB = bb + 4 * (size_t)p;
    PB[0] = fminf(LB[0], RB[0]);
    PB[1] = fminf(LB[1], RB[1]);
    PB[2] = fmaxf(LB[2], RB[2]);
    PB[3] = fmaxf(LB[3], RB[3]);
    float total = L[2] + R[2];
    P[2] = total;
    if (total > 0.0f) {
        P[0] = (L[0] * L[2] + R[0] * R[2]) / total;
        P[1] = (L[1] * L[2] + R[1] * R[2]) / total;
    } else {
        P[0] = 0.0f;
        P[1] = 0.0f;
    }
    float sx = PB[2] - PB[0], sy = PB[3] - PB[1];
    P[3] = fmaxf(sx, sy);
}

void orc_fill(int64_t n, const uint32_t *idx, const float *mass, const float *pos_xy,
              const uint32_t *left, const uint32_t *right, const uint32_t *parent,
              uint32_t offset, uint64_t order_seed, float *node_cm, float *node_aabb) {
    if (n <= 0) return;
    uint32_t *status = (uint32_t *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(uint32_t));
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) order[k] = (uint32_t)((k + offset) % n); /* lbvh.wgsl:171 */
    if (order_seed) {
        uint64_t st = order_seed;
        for (int64_t k = n - 1; k > 0; --k) {
            int64_t r = (int64_t)(splitmix64(&st) % (uint64_t)(k + 1));
            uint32_t tmp = order[k];
            order[k] = order[r];
            order[r] = tmp;
        }
    }
    for (int64_t k = 0; k < n; ++k) {
        uint32_t i = order[k];
        uint32_t cur = (uint32_t)(n - 1) + i;
        uint32_t b = idx[i];
        float px = pos_xy[2 * (size_t)b], py = pos_xy[2 * (size_t)b + 1];
        float *C = node_cm + 4 * (size_t)cur, *B = node_aabb + 4 * (size_t)cur;
        /* lbvh.wgsl:179-187 */
        C[0] = px;
        C[1] = py;
        C[2] = mass[b];
        C[3] = 0.0f;
        B[0] = px;
        B[1] = py;
        B[2] = px;
        B[3] = py;
        /* lbvh.wgsl:191-232 */
        while (cur != 0) {
            uint32_t p = parent[cur];
            if (p == cur) break;
            uint32_t prev = status[p]++;
            if (prev == 0) break;
            merge_node(p, left, right, node_cm, node_aabb);
            cur = p;
        }
    }
    free(order);
    free(status);
}

/* ------------------------------------------------------------------ walk */

void orc_walk(int64_t n, const uint32_t *idx, const float *node_cm, const uint32_t *left,
              const uint32_t *right, const uint32_t *roots, const float *mass,
              const float *pos_xy, const orc_params *p, int stack_size, float *accel_xy,
              int64_t *stats, const float *quad, float *margin, float *term_norm, int threads) {
    if (n <= 0) return;
    const float g = p->grav_constant, dt = p->delta_time, theta = p->bh_theta;
    const float mult = p->epsilon_multiplier, epsc = p->epsilon;
    const int ref_soft = p->softening_mode == ORC_SOFTEN_REFERENCE;
    const uint32_t leaf0 = (uint32_t)(n - 1);
    if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
#endif
    for (int64_t t = 0; t < n; ++t) {
        uint32_t stack[1024];
        int ssz = stack_size > 1024 ? 1024 : stack_size;
        uint32_t bi = idx[t]; /* barnes_hut.wgsl:56 */
        float p1x = pos_xy[2 * (size_t)bi], p1y = pos_xy[2 * (size_t)bi + 1];
        float m1 = mass[bi];
        float ax = 0.0f, ay = 0.0f;
        int64_t acc = 0, opened = 0, dropped = 0;
        float closest = INFINITY; /* smallest |length*inv - theta| / theta of an internal node */
        float l1 = 0.0f;          /* sum of the accepted terms' lengths */
        int sp = 0;
        stack[0] = roots ? roots[t] : 0u; /* barnes_hut.wgsl:67-70 */
        while (sp >= 0) {                  /* barnes_hut.wgsl:72-107 */
            uint32_t node = stack[sp];
            const float *C = node_cm + 4 * (size_t)node;
            sp -= 1;
            float rx = C[0] - p1x;
            float ry = C[1] - p1y;
            float d2 = rx * rx + ry * ry;
            int leaf = node >= leaf0;
            if (leaf && idx[node - leaf0] == bi) continue;
            float m2 = C[2];
            float eps;
            if (ref_soft) {
                float m_eff = fmaxf(m1, m2);
                eps = mult * sqrtf(g * m_eff * dt);
            } else {
                eps = epsc;
            }
            float inv = 1.0f / sqrtf(d2 + eps * eps);
            if (margin && !leaf) {
                float q = fabsf(C[3] * inv - theta) / theta;
                if (!(q >= closest)) closest = q;
            }
            if (leaf || C[3] * inv < theta) {
                float inv3 = inv * inv * inv;
                float gm = g * m2;
                if (quad) {
                    /* quadrupole term of the softened potential about the
                     * node's centre: -G inv^5 [(1.5 tr Q - 7.5 inv^2 r.Q r) r + 3 Q r] */
                    const float *Q = quad + 3 * (size_t)node;
                    float inv2 = inv * inv;
                    float inv5 = (inv2 * inv2) * inv;
                    float qrx = Q[0] * rx + Q[1] * ry, qry = Q[1] * rx + Q[2] * ry;
                    float rqr = rx * qrx + ry * qry;
                    float k = inv5 * (1.5f * (Q[0] + Q[2]) - 7.5f * inv2 * rqr);
                    ax = ax + (gm * rx * inv3 - g * (k * rx + 3.0f * inv5 * qrx));
                    ay = ay + (gm * ry * inv3 - g * (k * ry + 3.0f * inv5 * qry));
                } else {
                    ax = ax + gm * rx * inv3;
                    ay = ay + gm * ry * inv3;
                }
                if (term_norm) l1 += fabsf(gm * inv3) * sqrtf(d2);
                acc++;
            } else if (sp + 2 < ssz) {
                stack[++sp] = left[node];
                stack[++sp] = right[node];
                opened++;
            } else {
                dropped++;
            }
        }
        accel_xy[2 * (size_t)bi] = ax;
        accel_xy[2 * (size_t)bi + 1] = ay;
        if (stats) {
            stats[3 * (size_t)bi] = acc;
            stats[3 * (size_t)bi + 1] = opened;
            stats[3 * (size_t)bi + 2] = dropped;
        }
        if (margin) margin[bi] = closest;
        if (term_norm) term_norm[bi] = l1;
    }
}

int64_t orc_walk_lists(int64_t n, const uint32_t *idx, const float *node_cm, const uint32_t *left,
                       const uint32_t *right, const float *mass, const float *pos_xy,
                       const orc_params *p, int64_t *offsets, uint32_t *nodes, int64_t cap) {
    /* Same traversal and test as orc_walk (barnes_hut.wgsl:67-106). */
    const float g = p->grav_constant, dt = p->delta_time, theta = p->bh_theta;
    const float mult = p->epsilon_multiplier, epsc = p->epsilon;
    const int ref_soft = p->softening_mode == ORC_SOFTEN_REFERENCE;
    const uint32_t leaf0 = (uint32_t)(n - 1);
    int64_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && total > cap) return total;
        int64_t w = 0;
        for (int64_t t = 0; t < n; ++t) {
            if (pass == 0) offsets[t] = total;
            uint32_t stack[256];
            uint32_t bi = idx[t];
            float p1x = pos_xy[2 * (size_t)bi], p1y = pos_xy[2 * (size_t)bi + 1], m1 = mass[bi];
            int sp = 0;
            stack[0] = 0;
            while (sp >= 0) {
                uint32_t node = stack[sp--];
                const float *C = node_cm + 4 * (size_t)node;
                float rx = C[0] - p1x, ry = C[1] - p1y;
                float d2 = rx * rx + ry * ry;
                int leaf = node >= leaf0;
                if (leaf && idx[node - leaf0] == bi) continue;
                float eps = ref_soft ? mult * sqrtf(g * fmaxf(m1, C[2]) * dt) : epsc;
                float inv = 1.0f / sqrtf(d2 + eps * eps);
                if (leaf || C[3] * inv < theta) {
                    if (pass == 0)
                        total++;
                    else
                        nodes[w++] = node;
                } else if (sp + 2 < 256) {
                    stack[++sp] = left[node];
                    stack[++sp] = right[node];
                }
            }
        }
        if (pass == 0) offsets[n] = total;
    }
    return total;
}

void orc_kick_drift(int64_t n, const float *accel_xy, float dt, float *pos_xy, float *vel_xy) {
    for (int64_t i = 0; i < 2 * n; ++i) vel_xy[i] = vel_xy[i] + accel_xy[i] * dt; /* barnes_hut.wgsl:110 */
    for (int64_t i = 0; i < 2 * n; ++i) pos_xy[i] = pos_xy[i] + vel_xy[i] * dt;   /* barnes_hut.wgsl:120 */
}

/* ------------------------------------------------------------------ step */

typedef struct scratch {
    uint32_t *keys, *idx, *left, *right, *parent;
    float *cm, *bb, *acc;
} scratch;

static int scratch_alloc(scratch *s, int64_t n) {
    size_t nn = (size_t)(n > 1 ? n : 1), nodes = 2 * nn - 1;
    s->keys = (uint32_t *)malloc(4 * nn);
    s->idx = (uint32_t *)malloc(4 * nn);
    s->left = (uint32_t *)malloc(4 * nn);
    s->right = (uint32_t *)malloc(4 * nn);
    s->parent = (uint32_t *)malloc(4 * nodes);
    s->cm = (float *)malloc(16 * nodes);
    s->bb = (float *)malloc(16 * nodes);
    s->acc = (float *)malloc(8 * nn);
    return (s->keys && s->idx && s->left && s->right && s->parent && s->cm && s->bb && s->acc) ? 0 : -1;
}

static void scratch_free(scratch *s) {
    free(s->keys);
    free(s->idx);
    free(s->left);
    free(s->right);
    free(s->parent);
    free(s->cm);
    free(s->bb);
    free(s->acc);
}

static void one_step(int64_t n, const orc_params *p, const float *mass, float *pos, float *vel,
                     scratch *s, int64_t *stats, float *bbox_out, int threads) {
    float bbox[4];
    if (p->morton_mode == ORC_MORTON_BBOX30) {
        orc_bbox(n, pos, bbox);
        orc_morton_bbox30(n, pos, bbox, s->keys, s->idx);
    } else {
        orc_bbox(n, pos, bbox);
        orc_morton_ref(n, pos, s->keys, s->idx);
    }
    if (bbox_out) memcpy(bbox_out, bbox, sizeof(bbox));
    orc_sort_pairs_u32(n, s->keys, s->idx, 32);
    orc_build_u32(n, s->keys, s->left, s->right, s->parent);
    orc_fill(n, s->idx, mass, pos, s->left, s->right, s->parent, 0, 0, s->cm, s->bb);
    orc_walk(n, s->idx, s->cm, s->left, s->right, NULL, mass, pos, p, 256, s->acc, stats, NULL,
             NULL, NULL, threads);
    orc_kick_drift(n, s->acc, p->delta_time, pos, vel);
}

int orc_step(int64_t n, const orc_params *p, const float *mass, float *pos_xy, float *vel_xy,
             int nsteps, int threads) {
    if (n <= 0) return 0;
    scratch s;
    if (scratch_alloc(&s, n)) {
        scratch_free(&s);
        return -1;
    }
    for (int k = 0; k < nsteps; ++k) one_step(n, p, mass, pos_xy, vel_xy, &s, NULL, NULL, threads);
    scratch_free(&s);
    return 0;
}

int orc_step_debug(int64_t n, const orc_params *p, const float *mass, const float *pos_xy,
                   const float *vel_xy, uint32_t *keys, uint32_t *idx, uint32_t *left,
                   uint32_t *right, uint32_t *parent, float *node_cm, float *node_aabb,
                   float *accel_xy, float *pos_out, float *vel_out, int64_t *stats,
                   float *bbox4, int threads) {
    if (n <= 0) return 0;
    scratch s;
    if (scratch_alloc(&s, n)) {
        scratch_free(&s);
        return -1;
    }
    memcpy(pos_out, pos_xy, 8 * (size_t)n);
    memcpy(vel_out, vel_xy, 8 * (size_t)n);
    one_step(n, p, mass, pos_out, vel_out, &s, stats, bbox4, threads);
    /* re-derive the keys (the sort consumed them in place) */
    memcpy(keys, s.keys, 4 * (size_t)n);
    memcpy(idx, s.idx, 4 * (size_t)n);
    if (n > 1) {
        memcpy(left, s.left, 4 * (size_t)(n - 1));
        memcpy(right, s.right, 4 * (size_t)(n - 1));
    }
    memcpy(parent, s.parent, 4 * (size_t)(2 * n - 1));
    memcpy(node_cm, s.cm, 16 * (size_t)(2 * n - 1));
    memcpy(node_aabb, s.bb, 16 * (size_t)(2 * n - 1));
    memcpy(accel_xy, s.acc, 8 * (size_t)n);
    scratch_free(&s);
    return 0;
}

/* ------------------------------------------------------------ direct sum */

void orc_direct_sum_f64(int64_t n, const float *pos_xy, const float *mass, const orc_params *p,
                        const int64_t *targets, int64_t ntargets, double *acc_xy, int threads) {
    if (!targets) ntargets = n;
    const double g = p->grav_constant, dt = p->delta_time, mult = p->epsilon_multiplier;
    const double epsc = p->epsilon;
    const int ref_soft = p->softening_mode == ORC_SOFTEN_REFERENCE;
    if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
    for (int64_t k = 0; k < ntargets; ++k) {
        int64_t i = targets ? targets[k] : k;
        double xi = pos_xy[2 * i], yi = pos_xy[2 * i + 1], mi = mass[i];
        double ax = 0.0, ay = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double rx = (double)pos_xy[2 * j] - xi, ry = (double)pos_xy[2 * j + 1] - yi;
            double mj = mass[j];
            double eps = ref_soft ? mult * sqrt(g * (mi > mj ? mi : mj) * dt) : epsc;
            double s = rx * rx + ry * ry + eps * eps;
            double inv = 1.0 / sqrt(s);
            double f = g * mj * inv * inv * inv;
            ax += f * rx;
            ay += f * ry;
        }
        acc_xy[2 * k] = ax;
        acc_xy[2 * k + 1] = ay;
    }
}

/* ---------------------------------------------------------------- render */

/* exp(x), x <= 0, in binary32: x log2(e) = n + f (n = rint), 2^f by the
 * degree-7 Taylor polynomial of exp(f ln 2) in Horner form, times 2^n. */
float orc_render_exp(float x) {
    float t = x * 1.44269504f;
    float n = rintf(t);
    float f = t - n;
    float p = 1.5252734e-05f;
    p = p * f + 1.5403530e-04f;
    p = p * f + 1.3333558e-03f;
    p = p * f + 9.6181291e-03f;
    p = p * f + 5.5504109e-02f;
    p = p * f + 2.4022651e-01f;
    p = p * f + 6.9314718e-01f;
    p = p * f + 1.0f;
    return ldexpf(p, (int)n);
}

uint16_t orc_fixed40_to_f16(uint64_t v) {
    if (v == 0) return 0;
    int msb = 63 - clz64(v);
    int e = msb - 40;
    uint64_t m, rem, half;
    if (e >= -14) {
        int shift = msb - 10;
        m = v >> shift;
        rem = v & ((1ull << shift) - 1ull);
        half = 1ull << (shift - 1);
    } else {
        m = v >> 16;
        rem = v & 0xFFFFull;
        half = 0x8000ull;
        e = -15;
    }
    if (rem > half || (rem == half && (m & 1ull))) ++m;
    if (e == -15) return (uint16_t)m;
    if (m == 2048ull) {
        m = 1024ull;
        ++e;
    }
    if (e > 15) return 0x7C00u;
    return (uint16_t)(((uint32_t)(e + 15) << 10) | (uint32_t)(m & 0x3FFull));
}

void orc_splat(int64_t n, const float *pos_xy, const float *radius, const float *cam4,
               int32_t width, int32_t height, uint64_t *fixed) {
    /* density.wgsl:41-96, each fragment's value added exactly in units of 2^-40 */
    memset(fixed, 0, 8 * (size_t)width * (size_t)height);
    const float cx = cam4[0], cy = cam4[1], hx = cam4[2], hy = cam4[3];
    const float zoom = fmaxf(hx, hy);
    const float zoom_scale = 1.0f / (1.0f + 0.05f * zoom);
    const float sx = 2.0f / (float)width, sy = 2.0f / (float)height;
    const float W = (float)width, H = (float)height;
    for (int64_t i = 0; i < n; ++i) {
        const float px = pos_xy[2 * i], py = pos_xy[2 * i + 1], rm = radius[i];
        const float cnx = (px - cx) / hx, cny = (py - cy) / hy;
        const float rpx = 2.0f * rm;
        const float rnx = rpx * sx, rny = rpx * sy;
        if (!(fabsf(cnx) < 1e30f) || !(fabsf(cny) < 1e30f) || !(rnx > 0.0f) || !(rny > 0.0f) ||
            !(rnx < 1e30f))
            continue;
        const float base = rm > 1.0f ? 0.1f : 0.02f;
        const float fx0 = floorf((cnx - rnx + 1.0f) / sx - 0.5f) - 1.0f;
        const float fx1 = ceilf((cnx + rnx + 1.0f) / sx - 0.5f) + 1.0f;
        const float fy0 = floorf((1.0f - cny - rny) / sy - 0.5f) - 1.0f;
        const float fy1 = ceilf((1.0f - cny + rny) / sy - 0.5f) + 1.0f;
        const int x0 = (int)fminf(fmaxf(fx0, 0.0f), W), x1 = (int)fmaxf(fminf(fx1, W - 1.0f), -1.0f);
        const int y0 = (int)fminf(fmaxf(fy0, 0.0f), H), y1 = (int)fmaxf(fminf(fy1, H - 1.0f), -1.0f);
        for (int y = y0; y <= y1; ++y) {
            const float yn = 1.0f - ((float)y + 0.5f) * sy;
            const float v = (yn - cny) / rny;
            for (int x = x0; x <= x1; ++x) {
                const float xn = ((float)x + 0.5f) * sx - 1.0f;
                const float u = (xn - cnx) / rnx;
                const float r = sqrtf(u * u + v * v);
                if (r > 1.0f) continue;
                const float w = orc_render_exp(-10.0f * r * r);
                const float d = base * w * zoom_scale;
                fixed[(size_t)y * width + x] += (uint64_t)rint((double)d * 1099511627776.0);
            }
        }
    }
}

static double f16_value(uint32_t h) {
    uint32_t ex = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    double sgn = (h & 0x8000u) ? -1.0 : 1.0;
    if (ex == 0) return sgn * ldexp((double)mant, -24);
    if (ex == 31) return mant ? NAN : sgn * INFINITY;
    return sgn * ldexp((double)(mant | 0x400u), (int)ex - 25);
}

static uint32_t unorm8(double c) {
    double q = rint(c * 255.0);
    return (uint32_t)(q < 0.0 ? 0.0 : (q > 255.0 ? 255.0 : q));
}

static double mixd(double a, double b, double t) { return a * (1.0 - t) + b * t; }

void orc_tone_table(uint32_t *table) {
    /* tone_map.wgsl:66-84 in binary64 with the shader's binary32 constants */
    const double s1 = (double)0.33f, s2 = (double)0.66f, w3 = (double)0.34f;
    const double c1[3] = {(double)0.10f, (double)0.20f, (double)0.90f};
    const double c2[3] = {(double)0.20f, (double)0.90f, (double)0.90f};
    const double gamma = (double)(float)(1.0 / 2.2);
    for (uint32_t h = 0; h < 65536u; ++h) {
        double x = (double)8.0f * f16_value(h);
        double mapped = x / (1.0 + x);
        if (!(mapped == mapped)) mapped = x > 0.0 ? 1.0 : 0.0;
        mapped = fmin(fmax(mapped, 0.0), 1.0);
        double t = pow(mapped, gamma);
        uint32_t out = 255u << 24;
        for (int k = 0; k < 3; ++k) {
            double c;
            if (t < s1)
                c = mixd(0.0, c1[k], t / s1);
            else if (t < s2)
                c = mixd(c1[k], c2[k], (t - s1) / s1);
            else
                c = mixd(c2[k], 1.0, (t - s2) / w3);
            out |= unorm8(c) << (8 * k);
        }
        table[h] = out;
    }
}

void orc_render(int64_t n, const float *pos_xy, const float *radius, const float *cam4,
                int32_t width, int32_t height, uint16_t *d16, uint8_t *rgba) {
    size_t px = (size_t)width * (size_t)height;
    uint64_t *fixed = (uint64_t *)malloc(8 * px);
    uint32_t *table = (uint32_t *)malloc(4 * 65536);
    orc_splat(n, pos_xy, radius, cam4, width, height, fixed);
    orc_tone_table(table);
    for (int32_t y = 0; y < height; ++y)
        for (int32_t x = 0; x < width; ++x) {
            size_t texel = (size_t)(height - 1 - y) * width + x; /* tone_map.wgsl:58-70 uv flip */
            uint16_t h = orc_fixed40_to_f16(fixed[texel]);
            if (d16) d16[texel] = h;
            uint32_t c = table[h];
            uint8_t *o = rgba + 4 * ((size_t)y * width + x);
            o[0] = (uint8_t)c;
            o[1] = (uint8_t)(c >> 8);
            o[2] = (uint8_t)(c >> 16);
            o[3] = (uint8_t)(c >> 24);
        }
    free(fixed);
    free(table);
}

/* ------------------------------------------------------------ quadrupole */

void orc_node_quad(int64_t n, const uint32_t *idx, const float *mass, const float *pos_xy,
                   const uint32_t *left, const uint32_t *right, float *quad) {
    /* Central second mass moments {xx, xy, yy} of every node's leaf range,
     * in binary64: centre of mass first, then the moments about it, each a
     * sum over the range in sorted order.  Leaves: 0. */
    int64_t nodes = 2 * n - 1;
    if (n <= 0) return;
    int64_t *lo = (int64_t *)malloc(8 * (size_t)nodes), *hi = (int64_t *)malloc(8 * (size_t)nodes);
    int64_t *order = (int64_t *)malloc(8 * (size_t)nodes);
    for (int64_t t = 0; t < n; ++t) lo[n - 1 + t] = hi[n - 1 + t] = t;
    int64_t cnt = 0, sp = 0;
    int64_t *stack = (int64_t *)malloc(8 * (size_t)nodes);
    if (n > 1) stack[sp++] = 0;
    while (sp) { /* pre-order of internal nodes; reversed = children first */
        int64_t v = stack[--sp];
        order[cnt++] = v;
        if (left[v] < (uint32_t)(n - 1)) stack[sp++] = left[v];
        if (right[v] < (uint32_t)(n - 1)) stack[sp++] = right[v];
    }
    for (int64_t k = cnt - 1; k >= 0; --k) {
        int64_t v = order[k];
        lo[v] = lo[left[v]];
        hi[v] = hi[right[v]];
    }
    for (int64_t v = 0; v < nodes; ++v) {
        float *Q = quad + 3 * (size_t)v;
        Q[0] = Q[1] = Q[2] = 0.0f;
        if (v >= n - 1 && n > 1) continue;
        double m = 0.0, sx = 0.0, sy = 0.0;
        for (int64_t t = lo[v]; t <= hi[v]; ++t) {
            uint32_t b = idx[t];
            double mb = mass[b], xb = pos_xy[2 * (size_t)b], yb = pos_xy[2 * (size_t)b + 1];
            m += mb;
            sx += mb * xb;
            sy += mb * yb;
        }
        if (!(m > 0.0)) continue;
        double cx = sx / m, cy = sy / m, qxx = 0.0, qxy = 0.0, qyy = 0.0;
        for (int64_t t = lo[v]; t <= hi[v]; ++t) {
            uint32_t b = idx[t];
            double mb = mass[b], dx = pos_xy[2 * (size_t)b] - cx, dy = pos_xy[2 * (size_t)b + 1] - cy;
            qxx += mb * dx * dx;
            qxy += mb * dx * dy;
            qyy += mb * dy * dy;
        }
        Q[0] = (float)qxx;
        Q[1] = (float)qxy;
        Q[2] = (float)qyy;
    }
    free(stack);
    free(order);
    free(lo);
    free(hi);
}
