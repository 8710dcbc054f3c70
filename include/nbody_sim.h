/*
 * nbody_sim.h -- C ABI of the MI355X Barnes-Hut step (libnbody_sim.so).
 *
 * Drop-in boundary for the compute side of the reference's `Simulation`
 * class (src/simulation.ts:39-433) and its `Config` (src/config.ts:1-22).
 * Each entry point names the reference member it replaces.  Conventions:
 *   - every function returns NBODY_OK (0) or a negative NBODY_ERR_* code and
 *     never throws; nbody_last_error() gives the message;
 *   - host arrays are caller-owned; device memory is owned by the handle
 *     (as the reference's Simulation owns its GPUBuffers, simulation.ts:68-104);
 *   - pos/vel use the reference's host layout: float32 [x0,y0,x1,y1,...]
 *     (the vec2f buffers of simulation.ts:237-238, 376-380);
 *   - one handle is used by one host thread at a time; work is ordered on
 *     the handle's HIP stream (the reference orders it by queue submission,
 *     main.ts:38-41).
 */
#ifndef NBODY_SIM_H
#define NBODY_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBODY_OK 0
#define NBODY_ERR_INVALID_ARG (-1)
#define NBODY_ERR_HIP (-2)
#define NBODY_ERR_CAPACITY (-3)
#define NBODY_ERR_STACK_OVERFLOW (-4)
#define NBODY_ERR_STATE (-5)
#define NBODY_ERR_NO_DEVICE (-6)

/* Softening: NBODY_SOFTEN_REFERENCE is barnes_hut.wgsl:90-91,
 * eps = epsilon_multiplier * sqrt(G * max(m1, m2) * dt); NBODY_SOFTEN_CONSTANT
 * uses `epsilon` for every pair. */
#define NBODY_SOFTEN_REFERENCE 0
#define NBODY_SOFTEN_CONSTANT 1

/* Morton keys: NBODY_MORTON_REFERENCE is lbvh.wgsl:42-69 (fixed
 * [-1000,1000]^2 box, 16 bits per axis); NBODY_MORTON_BBOX30 quantises
 * 15 bits per axis over each step's square bounding box. */
#define NBODY_MORTON_REFERENCE 0
#define NBODY_MORTON_BBOX30 1

/* Walk arithmetic: FAST uses the hardware reciprocal square root and fused
 * multiply-adds; EXACT evaluates 1/sqrt and every product and sum as single
 * correctly rounded binary32 operations in the reference's order. */
#define NBODY_MATH_FAST 0
#define NBODY_MATH_EXACT 1

/* Expansion of an accepted node's pull: MONOPOLE is the reference's
 * G m r / (|r|^2 + eps^2)^{3/2} (barnes_hut.wgsl:92-97); QUADRUPOLE adds the
 * next term of the softened potential's Taylor series about the node's
 * centre of mass, from its second mass moments (no reference counterpart;
 * constant softening and the fast arithmetic only).  Same opening test. */
#define NBODY_EXPANSION_MONOPOLE 0
#define NBODY_EXPANSION_QUADRUPOLE 1

/* Replaces `Config` (config.ts:1-7) plus the `Metadata` uniform that
 * updateMetadataBuffer() packs (simulation.ts:212-225, lbvh.wgsl:3-10). */
typedef struct nbody_config {
    float grav_constant;      /* Config.gravConstant      (config.ts:2)  */
    float bh_theta;           /* Config.bhTheta           (config.ts:3)  */
    int32_t substeps;         /* Config.substeps          (config.ts:4)  */
    float delta_time;         /* Config.deltaTime         (config.ts:5)  */
    float epsilon_multiplier; /* Config.epsilonMultiplier (config.ts:6)  */
    float epsilon;            /* softening length for NBODY_SOFTEN_CONSTANT */
    int32_t softening_mode;   /* NBODY_SOFTEN_*  */
    int32_t morton_mode;      /* NBODY_MORTON_*  */
    int32_t math_mode;        /* NBODY_MATH_*    */
    uint32_t offset;          /* Metadata.offset (simulation.ts:222, lbvh.wgsl:171); the
                                 race-free fill gives the same tree for every value */
    int32_t expansion;        /* NBODY_EXPANSION_* (engine extension)                 */
} nbody_config;

typedef struct nbody_sim nbody_sim;

/* getDefaultConfig() (config.ts:9-22): G=1, theta=0.6, substeps=1,
 * dt=0.1/60, epsilonMultiplier=1, reference softening and keys. */
int nbody_default_config(nbody_config *out);

/* new Simulation(config, device) (simulation.ts:55-66).  Allocates device
 * storage for `capacity` bodies on HIP device `device` (grown on demand by
 * nbody_set_bodies).  *err (may be NULL) receives the status. */
nbody_sim *nbody_create(const nbody_config *cfg, int64_t capacity, int32_t device, int *err);

/* Ensemble handle: `num_sims` independent simulations with n_per_sim[s]
 * bodies each, stepped by one sort, one tree build and one walk per step.
 * Bodies are passed concatenated in sim order.  Keys are bbox30 codes per
 * sim prefixed by the sim id (num_sims <= 256).  No reference counterpart:
 * the reference steps one simulation (simulation.ts:39). */
nbody_sim *nbody_ensemble_create(const nbody_config *cfg, int32_t num_sims,
                                 const int64_t *n_per_sim, int32_t device, int *err);

int nbody_destroy(nbody_sim *sim);

/* updateMetadataBuffer() (simulation.ts:212-225): replace the parameters. */
int nbody_set_config(nbody_sim *sim, const nbody_config *cfg);
int nbody_get_config(const nbody_sim *sim, nbody_config *out);

/* setNumBodies(n) + updatePhysicsBuffers() (simulation.ts:227-231, 376-380):
 * upload n bodies.  For an ensemble handle n must equal the total body count. */
int nbody_set_bodies(nbody_sim *sim, int64_t n, const float *mass, const float *pos_xy,
                     const float *vel_xy);

/* getBuffers() (simulation.ts:427-429), read back to host.  Any pointer may
 * be NULL.  Synchronises the handle's stream. */
int nbody_get_bodies(nbody_sim *sim, float *mass, float *pos_xy, float *vel_xy);

/* getNumBodies() (simulation.ts:430-432). */
int64_t nbody_get_num_bodies(const nbody_sim *sim);

/* updateRadiusMultiplierBuffer() (simulation.ts:382-384) and the
 * radiusMultiplier member of getBuffers() (simulation.ts:17, 87, 427-429):
 * per-body size factors of the density splat (density.wgsl:57-66, bound by
 * renderer.ts:147-153).  The step never reads them.  nbody_set_bodies resets
 * them to 1, as setScenario fills them (simulation.ts:240-241). */
int nbody_set_radius_multiplier(nbody_sim *sim, const float *radius_multiplier);
int nbody_get_radius_multiplier(nbody_sim *sim, float *radius_multiplier);

/* Steps run since the bodies were last set (nbody_step, nbody_adjoint_forward;
 * restored by nbody_load_snapshot).  No reference counterpart. */
int nbody_get_step_count(const nbody_sim *sim, int64_t *steps);

/* State snapshot (no reference counterpart: the reference never reads its
 * buffers back).  nbody_save_snapshot writes directory `dir` (created if
 * needed): mass.npy f4[n], pos.npy f4[n,2], vel.npy f4[n,2],
 * radius_multiplier.npy f4[n], perm.npy u4[n] (sorted position -> body of the
 * last step; identity before any step) in NumPy .npy format, and meta.json
 * (n, simulation sizes, step count, config).  nbody_load_snapshot restores
 * config, bodies, radius multipliers, permutation and step count into a
 * handle with the same simulation layout; stepping on from a loaded snapshot
 * gives the same bits as stepping on from the saved handle. */
int nbody_save_snapshot(nbody_sim *sim, const char *dir);
int nbody_load_snapshot(nbody_sim *sim, const char *dir);

/* getCommands() + device.queue.submit() (simulation.ts:386-425, main.ts:38-41):
 * enqueue `nsteps` full steps (Morton encode, sort, tree build, fill, walk
 * with kick, drift) on the handle's stream.  Asynchronous. */
int nbody_step(nbody_sim *sim, int32_t nsteps);

/* Allocation-free step.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Enqueues `nsteps` plain
 * steps, exactly as nbody_step would for a handle with this config, default
 * tuning and the same bodies, of n <= NBODY_STATIC_MAX_BODIES bodies whose
 * mass (f32[n]), pos and vel (f32[2n], x0,y0,x1,y1,...) are caller-owned
 * DEVICE memory on the current HIP device; pos and vel are updated in place.
 * Every intermediate buffer lives in one fixed 16 MiB device array per GPU
 * (held from the library's load to the end of the process); the call makes
 * no hipMalloc/hipFree, no memset node and no host synchronisation, so it
 * can be captured into a HIP graph.  `hip_stream` is a hipStream_t of the
 * current device (NULL: the null stream).  Calls on one GPU share its array:
 * calls whose work overlaps on the device (other streams, other host
 * threads' streams) race and must be ordered by the caller.  Larger n:
 * NBODY_ERR_CAPACITY.  At these sizes the walk stack cannot overflow. */
#define NBODY_STATIC_MAX_BODIES 65536
int nbody_step_static(const nbody_config *cfg, int64_t n, const float *mass, float *pos_xy,
                      float *vel_xy, int32_t nsteps, void *hip_stream);

/* Wait for queued work; reports a walk stack overflow seen since the last call. */
int nbody_synchronize(nbody_sim *sim);

/* Run nbody_step's work on `hip_stream` (a hipStream_t; NULL = the handle's own). */
int nbody_set_stream(nbody_sim *sim, void *hip_stream);

/* Launch kernels one by one (0), or replay steps from a captured HIP graph
 * holding `steps_per_graph` (1..4096) consecutive steps: nbody_step(k) then
 * replays the graph k / steps_per_graph times and launches the remaining
 * k % steps_per_graph steps one by one (the reference encodes `substeps`
 * steps per command buffer, simulation.ts:393).  Results do not depend on it. */
int nbody_use_graph(nbody_sim *sim, int32_t steps_per_graph);

/* Capture the step graph now (graphs on, bodies set), so that the next
 * nbody_step replays it without capturing first; a no-op when the captured
 * graph is still valid.  nbody_step captures by itself when it needs to.
 * No counterpart in the reference, which encodes a new command buffer per
 * call (simulation.ts:386-425); for callers that keep capture out of a timed
 * or latency-critical loop. */
int nbody_prepare_graph(nbody_sim *sim);

/* Debug flags for later steps: NBODY_DEBUG_ACCEL records each body's
 * acceleration; NBODY_DEBUG_STATS counts accepted and opened nodes;
 * NBODY_DEBUG_FILL_CLIMB makes the tree fill complete its upper levels by
 * climbing through global arrival counters (the path it takes on its own
 * only when a level's node list is too long for workgroup memory);
 * NBODY_DEBUG_SORT_RADIX makes every sort take the radix passes (plain
 * handles otherwise sort up to 4096 bodies in one workgroup, and up to
 * 4194304 between the previous step's splitters; the permutation is the
 * same). */
#define NBODY_DEBUG_ACCEL 1
#define NBODY_DEBUG_STATS 2
#define NBODY_DEBUG_FILL_CLIMB 4
#define NBODY_DEBUG_SORT_RADIX 8
int nbody_set_debug(nbody_sim *sim, int32_t flags);

/* Intermediate buffers of the last step (reference buffers in comments). */
#define NBODY_BUF_KEYS 1         /* sorted Morton keys: u32[n] (u64[n] for ensembles); mortonCodes */
#define NBODY_BUF_BODY_INDICES 2 /* u32[n] sorted position -> body; bodyIndices */
#define NBODY_BUF_CHILDREN 3     /* u32[2*(n-1)] {left,right} per internal node; NodeData.left/right */
#define NBODY_BUF_PARENT 4       /* u32[2n-1]; NodeData.parent */
#define NBODY_BUF_NODE_CM 5      /* f32[4*(2n-1)] {com.x, com.y, total_mass, length} */
#define NBODY_BUF_NODE_AABB 6    /* f32[4*(2n-1)] {min.x, min.y, max.x, max.y} */
#define NBODY_BUF_ACCEL 7        /* f32[2n] per body (needs NBODY_DEBUG_ACCEL) */
#define NBODY_BUF_BBOX 8         /* f32[4*num_sims] bounding box of the last step */
#define NBODY_BUF_SIM_ROOTS 9    /* u32[num_sims] sub-tree root of each sim */
#define NBODY_BUF_WALK_COUNTS 10 /* u32[2n] per body {accepted, opened} of the last walk with
                                    NBODY_DEBUG_STATS (pair walks: walk_variant 2 and 4) */
#define NBODY_BUF_TONE_TABLE 11  /* u32[65536] RGBA8 (R in the low byte) of every fp16
                                    density value, after the first nbody_render */
#define NBODY_BUF_NODE_QUAD 12   /* f32[3*(2n-1)] {xx, xy, yy} central second mass moments of
                                    every node (quadrupole expansion; leaves 0) */
int nbody_debug_get(nbody_sim *sim, int32_t which, void *dst, int64_t bytes);

/* Walk counters since the last reset: {accepted, opened, overflowed}. */
int nbody_walk_stats(nbody_sim *sim, uint64_t *out3, int32_t reset);

/* Per-phase device time (HIP events on the handle's stream) while profiling
 * is enabled; phases: 0 bbox, 1 encode, 2 sort, 3 build, 4 fill, 5 walk+kick+drift.
 * ms_out[6] receives the sums since the last reset, *steps the step count. */
int nbody_set_profiling(nbody_sim *sim, int32_t enable);
int nbody_phase_times(nbody_sim *sim, float *ms_out, int64_t *steps, int32_t reset);

/* Launch tunables.  Block size, stack size, XCD order and bodies per lane
 * never change a result.  Variants 0 and 1 add every body's terms in the
 * reference's order and agree bit for bit; variants 2 and 4 take the same
 * decisions and add the same terms in their own fixed order, bit for bit
 * alike (fast mode only; in the exact mode they run variant 1).
 *   walk_block   walk workgroup size (64..1024, multiple of 64)
 *   walk_stack   walk stack entries per wavefront (8..256; the register-stack
 *                walks use one or two rows of 64: [64, 128]; the pair walk
 *                always has 63)
 *   walk_variant 0: stack in LDS; 1: stack in VGPR lanes; 2: pair walk, both
 *                children of an opened node tested per iteration, stack of
 *                (node, mask) in VGPR lanes (default); 4: the pair walk run
 *                by persistent wavefronts taking 64-body chunks from a
 *                device queue (ensembles: largest simulations first)
 *   xcd_order    order of the walk's workgroups over the 8 XCDs (each has its
 *                own L2): 0 hardware order (round-robin); 1 each XCD walks one
 *                contiguous run of Morton-ordered bodies; k > 1 runs of k
 *                workgroups dealt to the XCDs in turn (default 32)
 *   lane_bodies  bodies per lane of the VGPR-stack walk and the adjoint walk:
 *                1 (default), 2 or 4; a wavefront walks 64 * lane_bodies
 *                Morton-consecutive bodies together (the LDS and pair walks
 *                ignore it, walking one body per lane)
 *   radix_bits   digit width of the radix sort: 8, 10 (default) or 11 bits
 *                per pass (30-bit keys: 4, 3 or 3 passes)
 *   wave_bodies  bodies per wavefront of the pair walks (plain handles): 64,
 *                32, 16 or 8 Morton-consecutive bodies share one traversal;
 *                fewer give more, shorter traversals (small simulations,
 *                where the walk is bound by latency); 0 (default) = the
 *                widest that leaves at least 2048 wavefronts */
typedef struct nbody_tuning {
    int32_t walk_block;
    int32_t walk_stack;
    int32_t walk_variant;
    int32_t xcd_order;
    int32_t lane_bodies;
    int32_t radix_bits;
    int32_t wave_bodies;
} nbody_tuning;
int nbody_set_tuning(nbody_sim *sim, const nbody_tuning *t);
int nbody_get_tuning(const nbody_sim *sim, nbody_tuning *out);

/* Adjoint of the step (no reference counterpart: the reference only steps
 * forward).  nbody_adjoint_forward runs `nsteps` steps from the current state
 * like nbody_step, keeping a (pos, vel) checkpoint every `every` steps.
 * nbody_adjoint_backward then takes dL/dpos and dL/dvel of the final state
 * (host arrays, NULL = zero) and returns dL/dpos, dL/dvel and dL/dmass of
 * the initial state, with the tree topology and opening decisions of every
 * step frozen (each segment is recomputed from its checkpoint and every
 * step's tree rebuilt from its positions).  The final state is restored.
 * Plain handles, constant softening, n >= 2. */
int nbody_adjoint_forward(nbody_sim *sim, int32_t nsteps, int32_t every);
int nbody_adjoint_backward(nbody_sim *sim, const float *gpos_T, const float *gvel_T, float *gpos0,
                           float *gvel0, float *gmass);

/* Accuracy reference of the walk (no reference counterpart: the reference
 * never measures its approximation): binary64 softened direct sum over all
 * bodies of a plain handle, with the handle's G and softening, for
 * `ntargets` body indices `targets` (NULL: every body, ntargets ignored).
 * acc_xy receives double[2 * ntargets].  Synchronous; allocates scratch. */
int nbody_direct_accel_f64(nbody_sim *sim, const int64_t *targets, int64_t ntargets,
                           double *acc_xy);

/* The reference's Renderer (renderer.ts:22-329) drawing the simulation's
 * positions and radius multipliers: camera and viewport of its Metadata
 * uniform (renderer.ts:158-172, density.wgsl:3-10). */
typedef struct nbody_camera {
    float center_x, center_y; /* camCenter (renderer.ts:49)                       */
    float half_x, half_y;     /* camHalfSize (renderer.ts:50; x = y * aspect, 242-243) */
    int32_t width, height;    /* viewport and density texture size in pixels      */
} nbody_camera;

/* The Renderer's initial camera for a width x height canvas (renderer.ts:
 * 49-50, 240-243): centre (0, 0), half height 10, half width 10 * aspect. */
int nbody_default_camera(int32_t width, int32_t height, nbody_camera *out);

/* Renderer.getCommands() (renderer.ts:293-329) as compute kernels: every
 * body's quad of 2R x 2R pixels, R = 2 * radiusMultiplier, adds
 * base * exp(-10 r^2) * zoom_scale at each pixel centre with r = |uv| <= 1
 * (density.wgsl:41-96), summed exactly in fixed point and rounded once to
 * fp16 (the reference blends into r16float, rounding every add); then the
 * tone map (tone_map.wgsl:46-84: exposure 8, Reinhard, gamma 1/2.2, the
 * four-stop colour ramp) to 8-bit RGBA.  density_f16 (may be NULL) receives
 * the fp16 texture, uint16 bits [height][width], row 0 at the top (NDC y =
 * +1); rgba (may be NULL) the image, uint8 [height][width][4], row 0 at the
 * top of the canvas.  The full-screen pass flips rows, as the reference's
 * does: canvas row y shows texture row height-1-y.  Synchronous. */
int nbody_render(nbody_sim *sim, const nbody_camera *cam, uint16_t *density_f16, uint8_t *rgba);

/* 8-bit RGBA image -> PNG file (zlib-compressed, no filter). */
int nbody_write_png(const char *path, int32_t width, int32_t height, const uint8_t *rgba);

/* Message of the last failure on `sim` (NULL: the last failure of a call
 * without a usable handle).  Never NULL. */
const char *nbody_last_error(const nbody_sim *sim);

/* Initial conditions.  Scenarios 0-2 are setScenario()'s generators
 * (simulation.ts:243-362) with Math.random replaced by Philox-4x32-10;
 * 3-6 are the seeded workloads of the benchmark configurations. */
#define NBODY_SCENARIO_DEFAULT 0        /* simulation.ts:243-266 */
#define NBODY_SCENARIO_TWO_CLUSTERS 1   /* simulation.ts:267-306 */
#define NBODY_SCENARIO_PLANETARY_DISK 2 /* simulation.ts:308-361 */
#define NBODY_SCENARIO_UNIFORM_DISK 3   /* unit disk, v=0, m=1/n */
#define NBODY_SCENARIO_EXP_DISK 4       /* exponential disk, circular v + dispersion, lognormal m */
#define NBODY_SCENARIO_COLLISION 5      /* two exponential disks on a collision course */
#define NBODY_SCENARIO_PLUMMER 6        /* projected 2D Plummer, v=0, m=1/n */
typedef struct nbody_ic_params {
    float grav_constant; /* used by the circular-velocity scenarios */
    float scale;         /* disk scale length / Plummer radius (0 = scenario default) */
    float dispersion;    /* velocity dispersion as a fraction of v_circ (EXP_DISK) */
    float mass_sigma;    /* lognormal sigma of masses (EXP_DISK; 0 = equal masses) */
} nbody_ic_params;
int nbody_generate(int32_t scenario, int64_t n, uint64_t seed, const nbody_ic_params *params,
                   float *mass, float *pos_xy, float *vel_xy);

/* setScenario's radiusMultiplierData (simulation.ts:240-241, 321, 330, 339):
 * 1 for every body; planetary-disk: 12 for the central body, 5 for the two
 * planets. */
int nbody_generate_radius(int32_t scenario, int64_t n, float *radius_multiplier);

/* Philox-4x32-10 (Salmon et al., SC'11) used by nbody_generate. */
void nbody_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Library version string. */
const char *nbody_version(void);

#ifdef __cplusplus
}
#endif
#endif /* NBODY_SIM_H */
