/*
 * bh_oracle.h -- CPU restatement of the reference Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libnbody_sim.so) never links or calls it.
 *
 * Every function restates one phase of the reference pipeline
 * (src/simulation.ts:393-421 driving src/shaders/compute/lbvh.wgsl and
 * src/shaders/compute/barnes_hut.wgsl) in IEEE-754 single precision with no
 * FMA contraction (built with -ffp-contract=off), plus an FP64 direct sum.
 *
 * Node numbering follows lbvh.wgsl:94: internal nodes [0, n-2], leaves
 * [n-1, 2n-2].  Node arrays here are split by field:
 *   node_cm[4*k]   = {com.x, com.y, total_mass, length}   (lbvh.wgsl:12-22)
 *   node_aabb[4*k] = {aabb_min.x, aabb_min.y, aabb_max.x, aabb_max.y}
 *   left[k], right[k] for internal k in [0, n-2]; parent[k] for all 2n-1.
 */
#ifndef BH_ORACLE_H
#define BH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Softening and key modes (same values as include/nbody_sim.h). */
#define ORC_SOFTEN_REFERENCE 0 /* eps = mult*sqrt(G*max(m1,m2)*dt), barnes_hut.wgsl:90-91 */
#define ORC_SOFTEN_CONSTANT 1  /* eps = constant */
#define ORC_MORTON_REFERENCE 0 /* fixed [-1000,1000]^2 box, 16 bits/axis, lbvh.wgsl:42-69 */
#define ORC_MORTON_BBOX30 1    /* per-step bounding box, 15 bits/axis */

typedef struct orc_params {
    float grav_constant;      /* Metadata.grav_constant   (lbvh.wgsl:5)  */
    float delta_time;         /* Metadata.delta_time      (lbvh.wgsl:6)  */
    float epsilon_multiplier; /* Metadata.epsilon_multiplier (lbvh.wgsl:7) */
    float bh_theta;           /* Metadata.bh_theta        (lbvh.wgsl:8)  */
    float epsilon;            /* constant softening length (ORC_SOFTEN_CONSTANT) */
    int32_t softening_mode;
    int32_t morton_mode;
} orc_params;

/* lbvh.wgsl:42-69: 32-bit Morton code over the fixed box; idx[i] = i. */
void orc_morton_ref(int64_t n, const float *pos_xy, uint32_t *keys, uint32_t *idx);

/* Bounding box {minx, miny, maxx, maxy} of pos (NaN ignored, fminf/fmaxf). */
void orc_bbox(int64_t n, const float *pos_xy, float *bbox4);

/* 30-bit Morton code (15 bits/axis) over the square box anchored at the
 * bbox minimum with side max(extent); idx[i] = i. */
void orc_morton_bbox30(int64_t n, const float *pos_xy, const float *bbox4,
                       uint32_t *keys, uint32_t *idx);

/* Interleave two 16-bit coordinates: x -> even bits, y -> odd bits
 * (lbvh.wgsl:59-65). */
uint32_t orc_interleave16(uint32_t x, uint32_t y);

/* Stable LSD radix sort of (keys, vals) by the low `bits` bits of keys,
 * 8-bit digits.  Stands in for webgpu-radix-sort@1.0.7 (simulation.ts:132-140). */
void orc_sort_pairs_u32(int64_t n, uint32_t *keys, uint32_t *vals, int bits);
void orc_sort_pairs_u64(int64_t n, uint64_t *keys, uint32_t *vals, int bits);

/* lbvh.wgsl:72-83 delta() and lbvh.wgsl:85-165 build_lbvh_main, one loop
 * iteration per invocation.  parent[0] = 0xFFFFFFFF. */
void orc_build_u32(int64_t n, const uint32_t *keys, uint32_t *left, uint32_t *right,
                   uint32_t *parent);

/* Ensemble variant: 64-bit keys (sim id << 30 | morton30), tie-break on the
 * sim-local sorted index (i - sim_offset[sim]), keys whose sim differs are
 * compared by clz64.  sim_root[s] receives the node covering sim s. */
void orc_build_u64(int64_t n, const uint64_t *keys, const int64_t *sim_offset,
                   int32_t num_sims, uint32_t *left, uint32_t *right, uint32_t *parent,
                   uint32_t *sim_root);

/* lbvh.wgsl:168-233 fill_lbvh_main.  Invocations run one after another in
 * the order given by a seeded shuffle of the leaves (order_seed = 0: the
 * reference's rotation by `offset`, lbvh.wgsl:171).  Each invocation climbs
 * until it is the first arriver at a parent; the second arriver merges. */
void orc_fill(int64_t n, const uint32_t *idx, const float *mass, const float *pos_xy,
              const uint32_t *left, const uint32_t *right, const uint32_t *parent,
              uint32_t offset, uint64_t order_seed, float *node_cm, float *node_aabb);

/* barnes_hut.wgsl:48-111 bh_vel_step_main without the velocity write:
 * accel_xy[2*bi] for body bi = idx[t], t in [0,n).  roots (may be NULL):
 * per sorted position, the node the walk starts at (ensemble sub-tree
 * roots); NULL = node 0.  stats (may be NULL): per body {accepted, opened}.
 * Leaves are nodes [n-1, 2n-2]; the self leaf of sorted position t is n-1+t.
 * margin (may be NULL): per body, the smallest |length*inv - theta| / theta
 * over the internal nodes its walk tests (+inf if none): how close its
 * closest opening decision came to the other branch of barnes_hut.wgsl:94.
 * term_norm (may be NULL): per body, the sum of the lengths of the accepted
 * terms |G m r inv^3| (the scale of the sum's rounding error).
 * quad (may be NULL): per node {xx, xy, yy} central second mass moments
 * (orc_node_quad); each accepted node then adds the quadrupole term of the
 * softened potential, -G inv^5 [(1.5 tr Q - 7.5 inv^2 r.Q r) r + 3 Q r]
 * (an extension: the reference adds the monopole only).
 * Runs on `threads` OpenMP threads (results do not depend on it). */
void orc_walk(int64_t n, const uint32_t *idx, const float *node_cm, const uint32_t *left,
              const uint32_t *right, const uint32_t *roots, const float *mass,
              const float *pos_xy, const orc_params *p, int stack_size, float *accel_xy,
              int64_t *stats, const float *quad, float *margin, float *term_norm, int threads);

/* Central second mass moments {xx, xy, yy} (float[3 * (2n-1)]) of every
 * node's leaf range, summed in binary64 in sorted order about the range's
 * binary64 centre of mass; leaves 0 (a single body: the root is its leaf). */
void orc_node_quad(int64_t n, const uint32_t *idx, const float *mass, const float *pos_xy,
                   const uint32_t *left, const uint32_t *right, float *quad);

/* The nodes each sorted position's walk accepts (barnes_hut.wgsl:94), in
 * walk order, as CSR: offsets[n+1], nodes[offsets[t]..offsets[t+1]).
 * Returns the total count; when it exceeds `cap` nothing is written to
 * `nodes` and the caller retries with a larger buffer.  Used to freeze the
 * opening decisions of a step for the adjoint's FP64 reference. */
int64_t orc_walk_lists(int64_t n, const uint32_t *idx, const float *node_cm, const uint32_t *left,
                       const uint32_t *right, const float *mass, const float *pos_xy,
                       const orc_params *p, int64_t *offsets, uint32_t *nodes, int64_t cap);

/* barnes_hut.wgsl:110 then barnes_hut.wgsl:113-121: v += a*dt, x += v*dt. */
void orc_kick_drift(int64_t n, const float *accel_xy, float dt, float *pos_xy, float *vel_xy);

/* One full step (simulation.ts:393-421) on host arrays, `nsteps` times.
 * Scratch is allocated internally.  Returns 0 on success. */
int orc_step(int64_t n, const orc_params *p, const float *mass, float *pos_xy, float *vel_xy,
             int nsteps, int threads);

/* Per-phase debug run of one step: fills every intermediate. */
int orc_step_debug(int64_t n, const orc_params *p, const float *mass, const float *pos_xy,
                   const float *vel_xy, uint32_t *keys, uint32_t *idx, uint32_t *left,
                   uint32_t *right, uint32_t *parent, float *node_cm, float *node_aabb,
                   float *accel_xy, float *pos_out, float *vel_out, int64_t *stats,
                   float *bbox4, int threads);

/* FP64 softened direct sum over all sources for the given targets
 * (targets NULL = all n).  Softening per orc_params (reference mode uses
 * max(m_i, m_j)).  acc_xy[2*k] for target k. */
void orc_direct_sum_f64(int64_t n, const float *pos_xy, const float *mass,
                        const orc_params *p, const int64_t *targets, int64_t ntargets,
                        double *acc_xy, int threads);

/* ---- renderer (renderer.ts:293-329, density.wgsl, tone_map.wgsl) ----
 * cam4 = {center.x, center.y, half.x, half.y}; images are [height][width].
 * orc_splat: every body's quad of 2R x 2R pixels (R = 2 radius) adds
 * base * exp(-10 r^2) * zoom_scale (density.wgsl:77-96) at each pixel centre
 * with r <= 1, in fixed point (units of 2^-40), texture rows (row 0 = NDC
 * y +1).  orc_tone_table: RGBA8 (R in the low byte) of every fp16 density.
 * orc_render: fp16 density (texture rows) and RGBA (canvas rows: row y
 * shows texture row height-1-y, tone_map.wgsl:58-70). */
float orc_render_exp(float x);
uint16_t orc_fixed40_to_f16(uint64_t v);
void orc_splat(int64_t n, const float *pos_xy, const float *radius, const float *cam4,
               int32_t width, int32_t height, uint64_t *fixed);
void orc_tone_table(uint32_t *table);
void orc_render(int64_t n, const float *pos_xy, const float *radius, const float *cam4,
                int32_t width, int32_t height, uint16_t *d16, uint8_t *rgba);

#ifdef __cplusplus
}
#endif
#endif
