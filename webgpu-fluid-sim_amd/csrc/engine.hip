// engine.hip -- the step engine behind the C ABI (include/nbody_sim.h).
//
// Owns every device buffer of a simulation (the reference's SimBuffers,
// src/simulation.ts:8-18, 68-104) in one arena allocation, orders the six
// phases of a step on one HIP stream (the reference's getCommands(),
// simulation.ts:386-425), and can replay a step from a captured HIP graph.
//
// HBM layout (capacity C bodies, all 256-byte aligned):
//   mass f32[C], pos float2[C], vel float2[C]      -- by body index
//   rank u32[C] sorted position                    -- by body index (sort)
//   keys[2] u32|u64[C], vals[2] u32[C]             -- sort ping-pong
//   leaf float4[C] {x, y, mass, 0} in sorted order   -- the walk's own bodies
//   pcm float4[4C] pair records: children's {com, mass, length} + indices (walk)
//   paabb float4[2C] children's AABBs per internal node (fill)
//   root float4[2], child uint2[C], range uint2[C], parent u32[2C-1], status u32[C]
//   fill counters + top list + slots u32[C], sort scratch, bbox partials, debug accel float2[C],
//   counters
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <zlib.h>
#include <string>
#include <vector>

#include "../../include/nbody_sim.h"
#include "nbody_internal.h"

using namespace nbs;

namespace {

thread_local std::string g_last_error = "";

constexpr int kNumPhases = 6;

struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    template <typename T>
    T *take(int64_t count) {
        size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
        bytes = (bytes + 255) & ~(size_t)255;
        T *p = reinterpret_cast<T *>(base + used);
        used += bytes;
        return p;
    }
};

}  // namespace

struct nbody_sim {
    nbody_config cfg{};
    int device = 0;
    bool ensemble = false;
    int num_sims = 1;
    std::vector<int64_t> sim_off_host;  // [num_sims + 1]
    int64_t cap = 0, n = 0;
    float mass_absmax = 0.0f;  // largest |mass| uploaded (NaN if any is not a number)
    int key_bits = 32;
    int radix_bits = 10;  // sort digit width (nbody_tuning.radix_bits)

    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    Arena arena;

    float *mass = nullptr;
    float2 *pos = nullptr, *vel = nullptr;
    uint32_t *rank = nullptr;  // sorted position of each body (sort, for the build)
    void *keys[2] = {nullptr, nullptr};
    uint32_t *vals[2] = {nullptr, nullptr};
    TreeBufs tree{};
    uint32_t *sort_words = nullptr;
    SortScratch sort{};
    float4 *bbox_partial = nullptr, *bbox = nullptr;
    // Bounding-box accumulators: the walk folds the drifted positions into
    // them, the next step's encode reads them, the build zeroes them.
    // bbox_fresh: they hold the box of the current positions (false after
    // any change of the positions outside a step, or a tree-only step).
    uint32_t *bbox_slots = nullptr;
    bool bbox_fresh = false;
    int max_chunks = 1;    // bounding-box partials per sim (allocated)
    int sim_chunks = 1;    // bounding-box chunks of the largest sim (launched)
    int64_t *sim_off = nullptr;
    uint32_t *chunk_order = nullptr;  // ensembles: the queue walk's chunks, largest sims first
    uint32_t resident_waves = 0;      // wavefronts the GPU holds at once
    uint8_t *body_sim = nullptr;
    uint32_t *sim_root = nullptr;
    float2 *accel = nullptr;
    float *radius = nullptr;  // per-body radius multipliers (renderer input)
    double *quad_prefix = nullptr;  // leaf-moment prefix sums (quadrupole expansion)
    // renderer: fixed-point density, fp16 density, RGBA, tone table (one allocation)
    void *render_base = nullptr;
    int render_w = 0, render_h = 0;
    unsigned long long *render_dens = nullptr;
    uint16_t *render_d16 = nullptr;
    uint32_t *render_rgba = nullptr, *render_lut = nullptr;
    bool render_lut_ready = false;
    unsigned long long *stats = nullptr;
    uint2 *body_counts = nullptr;  // per-body walk counters (stats mode)
    uint32_t *overflow = nullptr;
    int result_buf = 0;  // which sort buffer holds the last sorted keys/values
    // bucket sort (plain handles): the splitters the last build wrote, each
    // body's bucket, and whether the splitters are those of the current
    // bodies' last step (else the sort takes the radix passes)
    uint32_t *split = nullptr;
    uint32_t *spill = nullptr;  // the bucket sort's oversized buckets: keys, then values
    bool split_valid = false;
    bool graph_cold_sort = false;  // the graph's first step sorts without splitters

    int32_t debug = 0;
    int walk_block = 256, walk_stack = 128;
    int walk_variant = 2;  // 0: LDS-stack walk, 1: register-stack walk, 2: pair walk (fast mode)
    int lane_bodies = 1;   // bodies per lane in the register-stack walks (1, 2, 4)
    int wave_bodies = 0;   // bodies per wavefront in the pair walks (64, 32, 16, 8; 0 = auto)
    int xcd_order = 32;  // walk workgroup order over the XCDs (nbody_tuning.xcd_order)

    bool use_graph = false;
    int graph_steps = 1;  // steps captured per graph
    int64_t steps_taken = 0;  // steps run since the bodies were set (snapshots)
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    bool graph_dirty = true;

    bool profiling = false;
    hipEvent_t ev[kNumPhases + 1] = {};
    float phase_ms[kNumPhases] = {};
    int64_t profiled_steps = 0;

    // adjoint state (allocated by nbody_adjoint_forward)
    void *adj_base = nullptr;
    int64_t adj_n = 0;
    int32_t adj_steps = 0, adj_every = 0, adj_cps = 0;
    float2 *adj_cp = nullptr;    // [cps][2][n] checkpoints (pos, vel)
    float2 *adj_seg = nullptr;   // [every][n] positions inside a segment
    float2 *adj_lx = nullptr, *adj_lv = nullptr, *adj_la = nullptr, *adj_end = nullptr;
    float *adj_lm = nullptr;
    float4 *adj_cot = nullptr;

    std::string last_error;
};

namespace {

int fail(nbody_sim *s, int code, const std::string &msg) {
    if (s) s->last_error = msg;
    g_last_error = msg;
    return code;
}

int hip_fail(nbody_sim *s, hipError_t e, const char *what) {
    return fail(s, NBODY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(s, call)                                       \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return hip_fail((s), e_, #call); \
    } while (0)

StepParams step_params(const nbody_config &c) {
    StepParams p;
    p.grav = c.grav_constant;
    p.dt = c.delta_time;
    p.eps_mult = c.epsilon_multiplier;
    p.theta = c.bh_theta;
    p.eps = c.epsilon;
    p.soft_mode = c.softening_mode;
    p.morton_mode = c.morton_mode;
    p.math_mode = c.math_mode;
    p.expansion = c.expansion;
    open_params(c.bh_theta, p.open_scale, p.open_bias);
    return p;
}

int validate_config(const nbody_config *c) {
    if (!c) return fail(nullptr, NBODY_ERR_INVALID_ARG, "config is NULL");
    if (c->softening_mode != NBODY_SOFTEN_REFERENCE && c->softening_mode != NBODY_SOFTEN_CONSTANT)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "unknown softening_mode");
    if (c->morton_mode != NBODY_MORTON_REFERENCE && c->morton_mode != NBODY_MORTON_BBOX30)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "unknown morton_mode");
    if (c->math_mode != NBODY_MATH_FAST && c->math_mode != NBODY_MATH_EXACT)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "unknown math_mode");
    if (c->substeps < 1) return fail(nullptr, NBODY_ERR_INVALID_ARG, "substeps must be >= 1");
    if (c->expansion != NBODY_EXPANSION_MONOPOLE && c->expansion != NBODY_EXPANSION_QUADRUPOLE)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "unknown expansion");
    if (c->expansion == NBODY_EXPANSION_QUADRUPOLE &&
        (c->softening_mode != NBODY_SOFTEN_CONSTANT || c->math_mode != NBODY_MATH_FAST))
        return fail(nullptr, NBODY_ERR_INVALID_ARG,
                    "the quadrupole expansion needs constant softening and the fast arithmetic");
    return NBODY_OK;
}

void release_graph(nbody_sim *s) {
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    s->graph_dirty = true;
}

// Carve every buffer for `cap` bodies out of `a` (a null base: sizes only).
// external_bodies: mass, pos and vel belong to the caller (not laid out).
void layout(nbody_sim *s, Arena &a, int64_t cap, bool external_bodies) {
    const size_t key_size = s->ensemble ? 8 : 4;
    const int64_t nodes = 2 * cap - 1;
    a.used = 0;
    if (!external_bodies) {
        s->mass = a.take<float>(cap);
        s->pos = a.take<float2>(cap);
        s->vel = a.take<float2>(cap);
    }
    s->rank = a.take<uint32_t>(cap);
    for (int b = 0; b < 2; ++b) {
        s->keys[b] = key_size == 8 ? (void *)a.take<uint64_t>(cap) : (void *)a.take<uint32_t>(cap);
        s->vals[b] = a.take<uint32_t>(cap);
    }
    s->tree.leaf = a.take<float4>(cap);
    s->tree.pcm = a.take<float4>(kPairStride * cap);
    s->tree.paabb = a.take<float4>(2 * cap);
    s->tree.root = a.take<float4>(4);
    s->tree.child = a.take<uint2>(cap);
    s->tree.range = a.take<uint2>(cap);
    s->tree.parent = a.take<uint32_t>(nodes);
    s->tree.status = a.take<uint32_t>(cap);
    s->tree.counts = a.take<uint32_t>(fill_count_words(cap));
    s->tree.lists = a.take<uint32_t>(fill_list_words(cap));
    s->tree.slot = a.take<uint32_t>(cap);
    s->sort_words = a.take<uint32_t>(sort_scratch_words(cap));
    s->bbox_partial = a.take<float4>((int64_t)s->max_chunks * s->num_sims);
    s->bbox = a.take<float4>(s->num_sims);
    s->bbox_slots = a.take<uint32_t>(4 * kBBoxSlots);
    s->sim_off = a.take<int64_t>(s->num_sims + 1);
    s->body_sim = s->ensemble ? a.take<uint8_t>(cap) : nullptr;
    s->sim_root = a.take<uint32_t>(s->num_sims);
    s->chunk_order = s->ensemble ? a.take<uint32_t>((cap + 63) / 64) : nullptr;
    s->accel = a.take<float2>(cap);
    const bool bucket_sort = !s->ensemble && !external_bodies && cap > kSortLdsMax;
    // the bucket sort's splitters
    s->split = bucket_sort ? a.take<uint32_t>(kMaxBuckets) : nullptr;
    s->spill = bucket_sort ? a.take<uint32_t>(2 * cap) : nullptr;
    s->radius = external_bodies ? nullptr : a.take<float>(cap);
    // the quadrupole pass's prefix sums (the allocation-free step is monopole only)
    s->quad_prefix = external_bodies ? nullptr : a.take<double>(quad_prefix_words(cap));
    s->stats = a.take<unsigned long long>(4);
    s->body_counts = a.take<uint2>(cap);
    s->overflow = a.take<uint32_t>(1);
    s->tree.fault = s->overflow;
}

// (Re)allocate every buffer for `cap` bodies.  Contents are not preserved.
int allocate(nbody_sim *s, int64_t cap) {
    release_graph(s);
    if (s->arena.base) {
        HIP_TRY(s, hipFree(s->arena.base));
        s->arena = Arena{};
    }
    cap = std::max<int64_t>(cap, 1);
    s->max_chunks = (int)std::max<int64_t>(1, (cap + kBBoxChunk - 1) / kBBoxChunk);
    const int64_t sort_words = sort_scratch_words(cap);
    // Size pass: lay out with a null base, then allocate the total.
    Arena &a = s->arena;
    layout(s, a, cap, false);
    a.size = a.used;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, a.size);
    if (e != hipSuccess) {
        s->cap = 0;
        return fail(s, NBODY_ERR_CAPACITY,
                    "hipMalloc of " + std::to_string(a.size) + " bytes failed: " +
                        hipGetErrorString(e));
    }
    a.base = static_cast<char *>(p);
    layout(s, a, cap, false);
    s->bbox_fresh = false;
    sort_bind_scratch(s->sort, s->sort_words, cap);
    HIP_TRY(s, hipMemsetAsync(s->sort_words, 0, 4 * (size_t)sort_words, s->stream));
    s->cap = cap;
    s->split_valid = false;
    HIP_TRY(s, hipMemsetAsync(s->stats, 0, 4 * sizeof(unsigned long long), s->stream));
    HIP_TRY(s, hipMemsetAsync(s->overflow, 0, sizeof(uint32_t), s->stream));
    HIP_TRY(s, hipMemsetAsync(s->tree.counts, 0, 4 * (size_t)fill_count_words(cap), s->stream));
    return NBODY_OK;
}

// True when a body's own leaf adds an exact zero to its fast-mode sum: the
// offset is exactly zero, and with constant softening 1/sqrt(eps^2), the
// factor inv^3 = eps^-3 (formed before the mass multiply) and |m| * eps^-3
// all stay finite (far from overflow), so the term m * inv^3 * 0 is a zero,
// never a NaN.
bool self_term_zero(const nbody_sim *s) {
    if (s->cfg.softening_mode != NBODY_SOFTEN_CONSTANT) return false;
    const double eps2 = (double)s->cfg.epsilon * (double)s->cfg.epsilon;
    const double m = (double)s->mass_absmax;
    if (!(eps2 >= 1e-30) || !(m <= 1e30)) return false;  // also false for NaN
    const double inv3 = std::pow(eps2, -1.5);
    // the quadrupole term also forms inv^5 (times zero moments for a leaf)
    if (s->cfg.expansion == NBODY_EXPANSION_QUADRUPOLE && !(std::pow(eps2, -2.5) <= 1e37))
        return false;
    return inv3 <= 1e37 && m * inv3 <= 1e30;
}

// The queue walk's chunk order for an ensemble: the chunks of 64 sorted
// positions of its simulations, largest simulation first (sim s occupies
// sorted positions [sim_off[s], sim_off[s+1]): its keys carry s as prefix);
// a chunk shared by two simulations goes with the larger one.
int upload_chunk_order(nbody_sim *s) {
    const int64_t total = s->sim_off_host[s->num_sims];
    const int64_t chunks = (total + 63) / 64;
    std::vector<int> sims(s->num_sims);
    for (int i = 0; i < s->num_sims; ++i) sims[i] = i;
    std::stable_sort(sims.begin(), sims.end(), [&](int x, int y) {
        return s->sim_off_host[x + 1] - s->sim_off_host[x] >
               s->sim_off_host[y + 1] - s->sim_off_host[y];
    });
    std::vector<uint8_t> seen((size_t)chunks, 0);
    std::vector<uint32_t> order;
    order.reserve((size_t)chunks);
    for (int sim : sims) {
        const int64_t lo = s->sim_off_host[sim], hi = s->sim_off_host[sim + 1];
        if (hi <= lo) continue;
        for (int64_t c = lo / 64; c <= (hi - 1) / 64; ++c)
            if (!seen[(size_t)c]) {
                seen[(size_t)c] = 1;
                order.push_back((uint32_t)c);
            }
    }
    HIP_TRY(s, hipMemcpyAsync(s->chunk_order, order.data(), 4 * order.size(),
                              hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

// Plain handles with bounding-box keys and two or more bodies carry the box
// from the walk of one step to the encode of the next (the accumulators);
// ensembles reduce their per-sim boxes with the bbox kernels every step.
bool uses_slots(const nbody_sim *s) {
    return !s->ensemble && s->cfg.morton_mode == NBODY_MORTON_BBOX30 && s->n > 1;
}

// Set the accumulators to the box of the current positions (the bbox
// kernels), unless they already hold it.  Not for graph capture.
void prime_bbox(nbody_sim *s) {
    if (!uses_slots(s) || s->bbox_fresh) return;
    launch_bbox(s->pos, s->sim_off, 1, s->max_chunks, s->bbox_partial, s->bbox, s->bbox_slots,
                s->stream);
    s->bbox_fresh = true;
}

// Which sort a plain step takes: one workgroup for n <= kSortLdsMax, the
// bucket sort between the last step's splitters when they are there, else
// (the first step after the bodies were set, ensembles, NBODY_DEBUG_SORT_RADIX)
// the radix passes.  Every path gives the same permutation.
enum class SortPath { kRadix, kSmall, kBuckets };
int64_t bucket_sort_max() { return (int64_t)kMaxBuckets << kBucketShiftMax; }
SortPath sort_path(const nbody_sim *s) {
    const int64_t n = s->n;
    if (s->ensemble || (s->debug & NBODY_DEBUG_SORT_RADIX) || n <= 1) return SortPath::kRadix;
    if (n <= kSortLdsMax) return SortPath::kSmall;
    if (s->split && s->split_valid && n <= bucket_sort_max()) return SortPath::kBuckets;
    return SortPath::kRadix;
}

// Enqueue one step: simulation.ts:393-421 (tree_only: phases 0-4, which
// rebuild the tree of the current positions without moving the bodies).
int enqueue_step(nbody_sim *s, bool tree_only = false) {
    const int64_t n = s->n;
    if (n <= 0) return NBODY_OK;
    hipStream_t st = s->stream;
    const StepParams p = step_params(s->cfg);
    const bool prof = s->profiling && !s->use_graph && !tree_only;
    auto mark = [&](int i) {
        if (prof) (void)hipEventRecord(s->ev[i], st);
    };

    mark(0);
    // phase 0: bounding box (bbox30 keys only): the ensembles' per-sim boxes,
    // or a plain handle's accumulators when the positions changed outside a
    // step (a graph capture happens after prime_bbox)
    const bool slots = uses_slots(s);
    if (slots)
        prime_bbox(s);
    else if (s->ensemble || s->cfg.morton_mode == NBODY_MORTON_BBOX30)
        launch_bbox(s->pos, s->sim_off, s->num_sims, s->ensemble ? s->sim_chunks : s->max_chunks,
                    s->bbox_partial, s->bbox,
                    nullptr, st);
    mark(1);
    // phase 1: Morton encode (lbvh.wgsl:42-69) and the sort's digit counts
    // (radix sort) or bucket counts (bucket sort)
    const int passes = n > 1 ? sort_passes(s->key_bits, s->radix_bits) : 0;
    const SortPath path = sort_path(s);
    uint32_t *keys0 = (uint32_t *)s->keys[0], *keys1 = (uint32_t *)s->keys[1];
    // the bucket sort: each tile of bodies grouped by bucket as it is encoded
    // (the tiles' bucket counts and starts in the look-back words, which this
    // path does not use otherwise)
    const int nb = path == SortPath::kBuckets ? sort_buckets(n) : 0;
    uint32_t *tcount = s->sort.lookback;
    uint32_t *tstart = s->sort.lookback + (int64_t)nb * encode_group_tiles(n);
    if (path == SortPath::kBuckets)
        launch_encode_group(s->pos, n, s->cfg.morton_mode, s->bbox,
                            slots ? s->bbox_slots : nullptr, keys1, s->vals[1], s->split, nb,
                            s->sort.hist, tcount, tstart, st);
    else
        launch_encode(s->pos, n, s->cfg.morton_mode, s->bbox, slots ? s->bbox_slots : nullptr,
                      s->body_sim, s->ensemble ? nullptr : keys0,
                      s->ensemble ? (uint64_t *)s->keys[0] : nullptr, s->vals[0], s->sort.hist,
                      path == SortPath::kRadix ? passes : 0, s->radix_bits, st);
    mark(2);
    // phase 2: stable sort (simulation.ts:400); it also writes each body's
    // sorted position
    const LeafGather lg{s->rank, s->overflow};
    if (s->ensemble)
        s->result_buf = launch_sort_u64((uint64_t *)s->keys[0], s->vals[0], (uint64_t *)s->keys[1],
                                        s->vals[1], n, s->key_bits, s->radix_bits, s->sort, lg, st);
    else if (path == SortPath::kSmall)
        s->result_buf = launch_sort_small(keys0, s->vals[0], keys1, s->vals[1], n, s->key_bits, lg, st);
    else if (path == SortPath::kBuckets) {
        if (launch_sort_gathered(keys1, s->vals[1], keys0, s->vals[0], s->spill, s->spill + s->cap,
                                 n, s->key_bits, nb, s->split, s->sort.hist, tcount, tstart, lg,
                                 st) != 0)
            return fail(s, NBODY_ERR_CAPACITY, "bucket sort: too many bodies");
        s->result_buf = 0;
    }
    else
        s->result_buf = launch_sort_u32(keys0, s->vals[0], keys1, s->vals[1], n, s->key_bits,
                                        s->radix_bits, s->sort, lg, st);
    const uint32_t *idx = s->vals[s->result_buf];
    mark(3);
    // phase 3: Karras build (lbvh.wgsl:85-165); it also places the sorted
    // leaf records and zeroes the sort's scratch for the next step
    const int64_t clear = sort_clear_words(n, s->key_bits, s->radix_bits, s->ensemble);
    // the build of a bucket-sortable size writes the next sort's splitters
    const bool splits = s->split && n > kSortLdsMax && n <= bucket_sort_max();
    const LeafPlace lpl{s->pos, s->mass, s->rank, splits ? s->split : nullptr, bucket_spacing(n)};
    if (s->ensemble)
        launch_build_u64((const uint64_t *)s->keys[s->result_buf], n, s->sim_off, s->tree, lpl,
                         s->sim_root, s->sort.hist, clear, nullptr, st);
    else
        launch_build_u32((const uint32_t *)s->keys[s->result_buf], n, s->tree, lpl, s->sort.hist,
                         clear, slots ? s->bbox_slots : nullptr, st);
    s->split_valid = splits;
    mark(4);
    // phase 4: bottom-up fill (lbvh.wgsl:168-233), then the second moments
    // of the quadrupole expansion
    // monopole pair records carry their children's opening keys (the
    // quadrupole pass writes moments into the same words)
    s->tree.pair_keys = p.expansion == NBODY_EXPANSION_MONOPOLE ? 1u : 0u;
    s->tree.open_scale = p.open_scale;
    s->tree.open_bias = p.open_bias;
    launch_fill(idx, s->mass, s->pos, n, s->tree, (s->debug & NBODY_DEBUG_FILL_CLIMB) != 0, st);
    if (s->cfg.expansion == NBODY_EXPANSION_QUADRUPOLE) launch_quad(s->tree, n, s->quad_prefix, st);
    mark(5);
    if (tree_only) {
        // the build has emptied the accumulators and no walk refills them
        s->bbox_fresh = false;
        return NBODY_OK;
    }
    // phase 5: walk + kick + drift (barnes_hut.wgsl:48-121)
    WalkArgs w{};
    w.leaf = s->tree.leaf;
    w.pcm = s->tree.pcm;
    w.root = s->tree.root;
    w.parent = s->tree.parent;
    w.child = s->tree.child;
    w.idx = idx;
    w.keys64 = s->ensemble ? (const uint64_t *)s->keys[s->result_buf] : nullptr;
    w.sim_root = s->sim_root;
    w.vel = s->vel;
    w.pos = s->pos;
    w.accel = (s->debug & NBODY_DEBUG_ACCEL) ? s->accel : nullptr;
    w.stats = s->stats;
    w.body_counts = (s->debug & NBODY_DEBUG_STATS) ? s->body_counts : nullptr;
    w.overflow = s->overflow;
    w.n = (uint32_t)n;
    w.p = p;
    w.self_zero = self_term_zero(s) ? 1u : 0u;
    w.queue = s->sort.counters + kWalkQueueWord;  // zeroed by the build
    w.chunk_order = s->chunk_order;
    w.resident = s->resident_waves;
    // the ensembles' queue order lists 64-body chunks: they keep 64 bodies per
    // wavefront; plain handles by default take the widest that leaves at
    // least 2048 wavefronts (two per SIMD), so a small simulation's
    // latency-bound traversals run side by side
    uint32_t wb = (uint32_t)s->wave_bodies;
    if (s->ensemble) {
        wb = 64u;
    } else if (wb == 0u) {
        wb = 64u;
        while (wb > 8u && (uint64_t)n / wb < 2048u) wb >>= 1;
    }
    w.wave_bodies = wb;
    w.bbox_slots = slots ? s->bbox_slots : nullptr;
    launch_walk(w, s->ensemble, (s->debug & NBODY_DEBUG_STATS) != 0, s->walk_block, s->walk_stack,
                s->xcd_order, s->walk_variant, s->lane_bodies, st);
    s->bbox_fresh = slots;  // the walk has folded the drifted positions in
    mark(6);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(s, e, "kernel launch");
    return NBODY_OK;
}

int accumulate_profile(nbody_sim *s) {
    HIP_TRY(s, hipEventSynchronize(s->ev[kNumPhases]));
    for (int i = 0; i < kNumPhases; ++i) {
        float ms = 0.0f;
        HIP_TRY(s, hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
        s->phase_ms[i] += ms;
    }
    s->profiled_steps += 1;
    return NBODY_OK;
}

int capture_graph(nbody_sim *s) {
    release_graph(s);
    // a graph whose first step cannot use splitters yet is captured again
    // once the splitters are there (nbody_step)
    const bool was_valid = s->split_valid;
    s->graph_cold_sort = sort_path(s) == SortPath::kRadix && s->split && !s->split_valid &&
                         s->n > kSortLdsMax && s->n <= bucket_sort_max();
    HIP_TRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = NBODY_OK;
    for (int k = 0; k < s->graph_steps && rc == NBODY_OK; ++k) rc = enqueue_step(s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    // nothing captured has run: the splitters are what they were
    s->split_valid = was_valid;
    if (rc != NBODY_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(s, e, "hipStreamEndCapture");
    s->graph = g;
    e = hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        release_graph(s);
        return hip_fail(s, e, "hipGraphInstantiate");
    }
    // move the graph's launch resources to the device now, not on the first replay
    e = hipGraphUpload(s->graph_exec, s->stream);
    if (e != hipSuccess) {
        release_graph(s);
        return hip_fail(s, e, "hipGraphUpload");
    }
    s->graph_dirty = false;
    return NBODY_OK;
}

// Capture the step graph unless the current one is still valid: the bbox
// priming runs eagerly first (captured steps expect it), and a graph whose
// first step had to sort without splitters is captured again once they exist.
int prepare_graph(nbody_sim *s) {
    prime_bbox(s);
    if (s->graph_dirty || !s->graph_exec || (s->graph_cold_sort && s->split_valid))
        return capture_graph(s);
    return NBODY_OK;
}

nbody_sim *create_common(const nbody_config *cfg, int32_t device, int *err, bool ensemble,
                         int num_sims, const int64_t *n_per_sim, int64_t capacity) {
    auto set_err = [&](int code) {
        if (err) *err = code;
    };
    int rc = validate_config(cfg);
    if (rc != NBODY_OK) {
        set_err(rc);
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err(fail(nullptr, NBODY_ERR_NO_DEVICE, "no HIP device available"));
        return nullptr;
    }
    if (device < 0 || device >= ndev) {
        set_err(fail(nullptr, NBODY_ERR_INVALID_ARG, "device index out of range"));
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_err(fail(nullptr, NBODY_ERR_HIP, "hipSetDevice failed"));
        return nullptr;
    }
    nbody_sim *s = new nbody_sim();
    s->cfg = *cfg;
    s->device = device;
    s->ensemble = ensemble;
    s->num_sims = num_sims;
    s->sim_off_host.assign(num_sims + 1, 0);
    if (ensemble) {
        int64_t largest = 1;
        for (int i = 0; i < num_sims; ++i) {
            s->sim_off_host[i + 1] = s->sim_off_host[i] + n_per_sim[i];
            largest = std::max<int64_t>(largest, n_per_sim[i]);
        }
        // the per-sim boxes' grid: the largest sim's chunks, not the whole
        // ensemble's (every other workgroup would find no bodies)
        s->sim_chunks = (int)((largest + kBBoxChunk - 1) / kBBoxChunk);
        capacity = s->sim_off_host[num_sims];
        int sim_bits = 0;
        while ((1 << sim_bits) < num_sims) ++sim_bits;
        s->key_bits = kSimShift + sim_bits;
    } else {
        s->key_bits = cfg->morton_mode == NBODY_MORTON_BBOX30 ? 30 : 32;
    }
    hipError_t e = hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_err(hip_fail(nullptr, e, "hipStreamCreate"));
        delete s;
        return nullptr;
    }
    s->stream = s->own_stream;
    for (auto &ev : s->ev) (void)hipEventCreate(&ev);
    rc = allocate(s, capacity);
    int cus = 0;
    if (rc == NBODY_OK &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess)
        s->resident_waves = (uint32_t)cus * 32u;  // 8 wavefronts per SIMD, 4 SIMDs per CU
    if (rc == NBODY_OK && ensemble) rc = upload_chunk_order(s);
    if (rc != NBODY_OK) {
        g_last_error = s->last_error;
        set_err(rc);
        nbody_destroy(s);
        return nullptr;
    }
    set_err(NBODY_OK);
    return s;
}

// Per-node buffers in the reference's numbering (internal [0, n-2], leaves
// [n-1, 2n-2]), reassembled on the host from the child-pair layout.
int debug_node_buffer(nbody_sim *s, int32_t which, void *dst, int64_t bytes) {
    const int64_t n = s->n;
    const int64_t nodes = std::max<int64_t>(2 * n - 1, 0);
    const int64_t need =
        nodes * (which == NBODY_BUF_PARENT ? 4 : (which == NBODY_BUF_NODE_QUAD ? 12 : 16));
    if (bytes < need)
        return fail(s, NBODY_ERR_INVALID_ARG,
                    "destination too small: need " + std::to_string(need) + " bytes");
    if (n <= 0) return NBODY_OK;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    std::vector<uint32_t> parent((size_t)nodes);
    HIP_TRY(s, hipMemcpy(parent.data(), s->tree.parent, 4 * (size_t)nodes, hipMemcpyDeviceToHost));
    if (which == NBODY_BUF_PARENT) {
        for (auto &v : parent)
            if (v != kNoNode) v &= kParentMask;
        std::memcpy(dst, parent.data(), 4 * (size_t)nodes);
        return NBODY_OK;
    }
    if (which == NBODY_BUF_NODE_QUAD) {
        // the monopole expansion computes no moments (its pair records' padding
        // holds opening keys)
        if (s->cfg.expansion != NBODY_EXPANSION_QUADRUPOLE) {
            std::memset(dst, 0, (size_t)need);
            return NBODY_OK;
        }
        // moments of child `side` of p: pcm[4p+2].zw, pcm[4p+3].x (left),
        // pcm[4p+3].yzw (right); the root's in root[2]
        std::vector<float> pr(16 * (size_t)std::max<int64_t>(n - 1, 1));
        float4 r2;
        if (n > 1)
            HIP_TRY(s, hipMemcpy(pr.data(), s->tree.pcm, 64 * (size_t)(n - 1),
                                 hipMemcpyDeviceToHost));
        HIP_TRY(s, hipMemcpy(&r2, s->tree.root + 2, 16, hipMemcpyDeviceToHost));
        float *out = static_cast<float *>(dst);
        out[0] = r2.x;
        out[1] = r2.y;
        out[2] = r2.z;
        for (int64_t k = 1; k < nodes; ++k) {
            const uint32_t pw = parent[(size_t)k];
            const float *rec = &pr[16 * (size_t)(pw & kParentMask)];
            const float *q = (pw >> 31) ? rec + 13 : rec + 10;
            if (k >= n - 1) q = nullptr;  // leaves have no moments
            for (int c = 0; c < 3; ++c) out[3 * k + c] = q ? q[c] : 0.0f;
        }
        return NBODY_OK;
    }
    const bool cm = which == NBODY_BUF_NODE_CM;
    const size_t per = cm ? kPairStride : 2;  // float4s per pair
    std::vector<float4> pair(per * (size_t)std::max<int64_t>(n - 1, 1)), root(2);
    const float4 *src = cm ? s->tree.pcm : s->tree.paabb;
    if (n > 1)
        HIP_TRY(s, hipMemcpy(pair.data(), src, 16 * per * (size_t)(n - 1),
                             hipMemcpyDeviceToHost));
    HIP_TRY(s, hipMemcpy(root.data(), s->tree.root, 32, hipMemcpyDeviceToHost));
    float4 *out = static_cast<float4 *>(dst);
    out[0] = which == NBODY_BUF_NODE_CM ? root[0] : root[1];
    for (int64_t k = 1; k < nodes; ++k) {
        const uint32_t pw = parent[(size_t)k];
        out[k] = pair[cm ? pcm_slot(pw) : paabb_slot(pw)];
    }
    // leaves carry kLeafLength on the device; the reference's leaf length is 0
    if (which == NBODY_BUF_NODE_CM)
        for (int64_t k = n - 1; k < nodes; ++k) out[k].w = 0.0f;
    return NBODY_OK;
}

}  // namespace

extern "C" {

int nbody_default_config(nbody_config *out) {
    if (!out) return fail(nullptr, NBODY_ERR_INVALID_ARG, "out is NULL");
    // config.ts:9-22
    std::memset(out, 0, sizeof(*out));
    out->grav_constant = 1.0f;
    out->bh_theta = 0.6f;
    out->substeps = 1;
    out->delta_time = (float)(0.1 * 1.0 / (60.0 * 1));
    out->epsilon_multiplier = 1.0f;
    out->epsilon = 1e-3f;
    out->softening_mode = NBODY_SOFTEN_REFERENCE;
    out->morton_mode = NBODY_MORTON_REFERENCE;
    out->math_mode = NBODY_MATH_FAST;
    out->offset = 0;
    return NBODY_OK;
}

nbody_sim *nbody_create(const nbody_config *cfg, int64_t capacity, int32_t device, int *err) {
    if (capacity < 0 || capacity >= (int64_t(1) << 30)) {
        int rc = fail(nullptr, NBODY_ERR_INVALID_ARG, "capacity must be in [0, 2^30)");
        if (err) *err = rc;
        return nullptr;
    }
    return create_common(cfg, device, err, false, 1, nullptr, capacity);
}

nbody_sim *nbody_ensemble_create(const nbody_config *cfg, int32_t num_sims,
                                 const int64_t *n_per_sim, int32_t device, int *err) {
    auto bad = [&](const char *msg) -> nbody_sim * {
        int rc = fail(nullptr, NBODY_ERR_INVALID_ARG, msg);
        if (err) *err = rc;
        return nullptr;
    };
    if (num_sims < 1 || num_sims > kMaxSims) return bad("num_sims must be in [1, 256]");
    if (!n_per_sim) return bad("n_per_sim is NULL");
    int64_t total = 0;
    for (int i = 0; i < num_sims; ++i) {
        if (n_per_sim[i] < 1) return bad("every sim needs at least one body");
        total += n_per_sim[i];
    }
    if (total >= (int64_t(1) << 30)) return bad("total bodies must be < 2^30");
    return create_common(cfg, device, err, true, num_sims, n_per_sim, total);
}

int nbody_destroy(nbody_sim *s) {
    if (!s) return NBODY_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    release_graph(s);
    if (s->arena.base) (void)hipFree(s->arena.base);
    if (s->adj_base) (void)hipFree(s->adj_base);
    if (s->render_base) (void)hipFree(s->render_base);
    for (auto &ev : s->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    delete s;
    return NBODY_OK;
}

int nbody_set_config(nbody_sim *s, const nbody_config *cfg) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    int rc = validate_config(cfg);
    if (rc != NBODY_OK) return fail(s, rc, g_last_error);
    if (!s->ensemble && cfg->morton_mode != s->cfg.morton_mode) {
        s->key_bits = cfg->morton_mode == NBODY_MORTON_BBOX30 ? 30 : 32;
    }
    if (s->ensemble && cfg->morton_mode != NBODY_MORTON_BBOX30 &&
        cfg->morton_mode != NBODY_MORTON_REFERENCE)
        return fail(s, NBODY_ERR_INVALID_ARG, "unknown morton_mode");
    s->cfg = *cfg;
    s->graph_dirty = true;
    s->split_valid = false;  // the keys may change form
    s->bbox_fresh = false;  // the accumulators are kept only in the bbox30 mode
    return NBODY_OK;
}

int nbody_get_config(const nbody_sim *s, nbody_config *out) {
    if (!s || !out) return fail(nullptr, NBODY_ERR_INVALID_ARG, "NULL argument");
    *out = s->cfg;
    return NBODY_OK;
}

int nbody_set_bodies(nbody_sim *s, int64_t n, const float *mass, const float *pos_xy,
                     const float *vel_xy) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (n < 0 || n >= (int64_t(1) << 30)) return fail(s, NBODY_ERR_INVALID_ARG, "n out of range");
    if (n > 0 && (!mass || !pos_xy || !vel_xy))
        return fail(s, NBODY_ERR_INVALID_ARG, "mass/pos/vel must be non-NULL");
    if (s->ensemble && n != s->sim_off_host[s->num_sims])
        return fail(s, NBODY_ERR_INVALID_ARG, "ensemble body count must equal the sum of n_per_sim");
    HIP_TRY(s, hipSetDevice(s->device));
    if (n > s->cap) {
        HIP_TRY(s, hipStreamSynchronize(s->stream));
        int rc = allocate(s, n);
        if (rc != NBODY_OK) return rc;
    }
    s->n = n;
    if (!s->ensemble) {
        s->sim_off_host[0] = 0;
        s->sim_off_host[1] = n;
    }
    float mmax = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        const float a = std::fabs(mass[i]);
        if (!(a <= mmax)) mmax = a;  // a NaN sticks
        if (mmax != mmax) break;
    }
    s->mass_absmax = mmax;
    if (n > 0) {
        HIP_TRY(s, hipMemcpyAsync(s->mass, mass, 4 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(s, hipMemcpyAsync(s->pos, pos_xy, 8 * (size_t)n, hipMemcpyHostToDevice, s->stream));
        s->bbox_fresh = false;
        HIP_TRY(s, hipMemcpyAsync(s->vel, vel_xy, 8 * (size_t)n, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(s, hipMemcpyAsync(s->sim_off, s->sim_off_host.data(), 8 * (size_t)(s->num_sims + 1),
                              hipMemcpyHostToDevice, s->stream));
    if (n > 0)  // setScenario's radiusMultiplierData.fill(1.0) (simulation.ts:240-241)
        launch_fill_f32(s->radius, 1.0f, n, s->stream);
    s->steps_taken = 0;
    s->split_valid = false;  // splitters of other bodies
    if (s->ensemble) {
        std::vector<uint8_t> bs((size_t)n);
        for (int i = 0; i < s->num_sims; ++i)
            std::fill(bs.begin() + s->sim_off_host[i], bs.begin() + s->sim_off_host[i + 1],
                      (uint8_t)i);
        HIP_TRY(s, hipMemcpyAsync(s->body_sim, bs.data(), (size_t)n, hipMemcpyHostToDevice,
                                  s->stream));
        HIP_TRY(s, hipStreamSynchronize(s->stream));  // bs goes out of scope
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    s->graph_dirty = true;
    return NBODY_OK;
}

int nbody_get_bodies(nbody_sim *s, float *mass, float *pos_xy, float *vel_xy) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->n;
    if (n > 0) {
        if (mass) HIP_TRY(s, hipMemcpyAsync(mass, s->mass, 4 * n, hipMemcpyDeviceToHost, s->stream));
        if (pos_xy) HIP_TRY(s, hipMemcpyAsync(pos_xy, s->pos, 8 * n, hipMemcpyDeviceToHost, s->stream));
        if (vel_xy) HIP_TRY(s, hipMemcpyAsync(vel_xy, s->vel, 8 * n, hipMemcpyDeviceToHost, s->stream));
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

int64_t nbody_get_num_bodies(const nbody_sim *s) { return s ? s->n : -1; }

int nbody_step(nbody_sim *s, int32_t nsteps) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (nsteps < 0) return fail(s, NBODY_ERR_INVALID_ARG, "nsteps must be >= 0");
    HIP_TRY(s, hipSetDevice(s->device));
    if (s->n <= 0 || nsteps == 0) return NBODY_OK;
    s->steps_taken += nsteps;
    if (s->use_graph) {
        const int k = s->graph_steps;
        if (nsteps >= k) {
            int rc = prepare_graph(s);
            if (rc != NBODY_OK) return rc;
        }
        for (int i = 0; i < nsteps / k; ++i) {
            HIP_TRY(s, hipGraphLaunch(s->graph_exec, s->stream));
            // every replay's builds write splitters for the current bodies
            s->split_valid = s->split && s->n > kSortLdsMax && s->n <= bucket_sort_max();
        }
        for (int i = 0; i < nsteps % k; ++i) {  // the rest, launched one by one
            int rc = enqueue_step(s);
            if (rc != NBODY_OK) return rc;
        }
        return NBODY_OK;
    }
    for (int i = 0; i < nsteps; ++i) {
        int rc = enqueue_step(s);
        if (rc != NBODY_OK) return rc;
        if (s->profiling) {
            rc = accumulate_profile(s);
            if (rc != NBODY_OK) return rc;
        }
    }
    return NBODY_OK;
}

int nbody_prepare_graph(nbody_sim *s) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    HIP_TRY(s, hipSetDevice(s->device));
    if (!s->use_graph || s->n <= 0) return NBODY_OK;
    return prepare_graph(s);
}

int nbody_synchronize(nbody_sim *s) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    HIP_TRY(s, hipSetDevice(s->device));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    uint32_t of = 0;
    HIP_TRY(s, hipMemcpy(&of, s->overflow, sizeof(of), hipMemcpyDeviceToHost));
    if (of) {
        HIP_TRY(s, hipMemset(s->overflow, 0, sizeof(uint32_t)));
        if (of & kFaultWait)
            return fail(s, NBODY_ERR_STATE,
                        "a workgroup's wait on another gave up (sort look-back or tree fill): "
                        "the step's results are not valid");
        return fail(s, NBODY_ERR_STACK_OVERFLOW,
                    "tree walk stack overflow: nodes were dropped (raise walk_stack)");
    }
    return NBODY_OK;
}

int nbody_set_stream(nbody_sim *s, void *hip_stream) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    HIP_TRY(s, hipSetDevice(s->device));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    s->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : s->own_stream;
    s->graph_dirty = true;
    return NBODY_OK;
}

int nbody_use_graph(nbody_sim *s, int32_t steps_per_graph) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (steps_per_graph < 0 || steps_per_graph > 4096)
        return fail(s, NBODY_ERR_INVALID_ARG, "steps_per_graph must be in [0, 4096]");
    s->use_graph = steps_per_graph != 0;
    s->graph_steps = std::max(1, (int)steps_per_graph);
    s->graph_dirty = true;
    return NBODY_OK;
}

int nbody_set_debug(nbody_sim *s, int32_t flags) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (flags & ~(NBODY_DEBUG_ACCEL | NBODY_DEBUG_STATS | NBODY_DEBUG_FILL_CLIMB |
                  NBODY_DEBUG_SORT_RADIX))
        return fail(s, NBODY_ERR_INVALID_ARG, "unknown debug flag");
    s->debug = flags;
    s->graph_dirty = true;
    return NBODY_OK;
}

int nbody_debug_get(nbody_sim *s, int32_t which, void *dst, int64_t bytes) {
    if (!s || !dst) return fail(s, NBODY_ERR_INVALID_ARG, "NULL argument");
    HIP_TRY(s, hipSetDevice(s->device));
    const int64_t n = s->n;
    const void *src = nullptr;
    int64_t need = 0;
    switch (which) {
        case NBODY_BUF_KEYS:
            src = s->keys[s->result_buf];
            need = n * (s->ensemble ? 8 : 4);
            break;
        case NBODY_BUF_BODY_INDICES:
            src = s->vals[s->result_buf];
            need = n * 4;
            break;
        case NBODY_BUF_CHILDREN:
            src = s->tree.child;
            need = std::max<int64_t>(n - 1, 0) * 8;
            break;
        case NBODY_BUF_PARENT:
        case NBODY_BUF_NODE_CM:
        case NBODY_BUF_NODE_AABB:
        case NBODY_BUF_NODE_QUAD:
            return debug_node_buffer(s, which, dst, bytes);
        case NBODY_BUF_ACCEL:
            src = s->accel;
            need = n * 8;
            break;
        case NBODY_BUF_BBOX:
            src = s->bbox;
            need = (int64_t)s->num_sims * 16;
            break;
        case NBODY_BUF_SIM_ROOTS:
            src = s->sim_root;
            need = (int64_t)s->num_sims * 4;
            break;
        case NBODY_BUF_WALK_COUNTS:
            src = s->body_counts;
            need = n * 8;
            break;
        case NBODY_BUF_TONE_TABLE:
            if (!s->render_lut_ready)
                return fail(s, NBODY_ERR_STATE, "no tone table before the first nbody_render");
            src = s->render_lut;
            need = 65536 * 4;
            break;
        default:
            return fail(s, NBODY_ERR_INVALID_ARG, "unknown buffer id");
    }
    if (bytes < need)
        return fail(s, NBODY_ERR_INVALID_ARG,
                    "destination too small: need " + std::to_string(need) + " bytes");
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (need > 0) HIP_TRY(s, hipMemcpy(dst, src, (size_t)need, hipMemcpyDeviceToHost));
    return NBODY_OK;
}

int nbody_walk_stats(nbody_sim *s, uint64_t *out3, int32_t reset) {
    if (!s || !out3) return fail(s, NBODY_ERR_INVALID_ARG, "NULL argument");
    HIP_TRY(s, hipSetDevice(s->device));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    unsigned long long h[4] = {0, 0, 0, 0};
    HIP_TRY(s, hipMemcpy(h, s->stats, sizeof(h), hipMemcpyDeviceToHost));
    out3[0] = h[0];
    out3[1] = h[1];
    out3[2] = h[2];
    if (reset) HIP_TRY(s, hipMemset(s->stats, 0, sizeof(h)));
    return NBODY_OK;
}

int nbody_set_profiling(nbody_sim *s, int32_t enable) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    s->profiling = enable != 0;
    return NBODY_OK;
}

int nbody_phase_times(nbody_sim *s, float *ms_out, int64_t *steps, int32_t reset) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (ms_out)
        for (int i = 0; i < kNumPhases; ++i) ms_out[i] = s->phase_ms[i];
    if (steps) *steps = s->profiled_steps;
    if (reset) {
        for (float &v : s->phase_ms) v = 0.0f;
        s->profiled_steps = 0;
    }
    return NBODY_OK;
}

int nbody_set_tuning(nbody_sim *s, const nbody_tuning *t) {
    if (!s || !t) return fail(s, NBODY_ERR_INVALID_ARG, "NULL argument");
    if (t->walk_block < 64 || t->walk_block > 1024 || t->walk_block % 64)
        return fail(s, NBODY_ERR_INVALID_ARG, "walk_block must be a multiple of 64 in [64, 1024]");
    if (t->walk_stack < 8 || t->walk_stack > 256)
        return fail(s, NBODY_ERR_INVALID_ARG, "walk_stack must be in [8, 256]");
    if (t->walk_variant < 0 || t->walk_variant > 4 || t->walk_variant == 3)
        return fail(s, NBODY_ERR_INVALID_ARG, "walk_variant must be 0, 1, 2 or 4");
    if (t->lane_bodies != 1 && t->lane_bodies != 2 && t->lane_bodies != 4)
        return fail(s, NBODY_ERR_INVALID_ARG, "lane_bodies must be 1, 2 or 4");
    if (t->wave_bodies != 0 && t->wave_bodies != 64 && t->wave_bodies != 32 &&
        t->wave_bodies != 16 && t->wave_bodies != 8)
        return fail(s, NBODY_ERR_INVALID_ARG, "wave_bodies must be 0 (auto), 64, 32, 16 or 8");
    if (t->xcd_order < 0 || t->xcd_order > 65536)
        return fail(s, NBODY_ERR_INVALID_ARG, "xcd_order must be in [0, 65536]");
    if (t->radix_bits != 8 && t->radix_bits != 10 && t->radix_bits != 11)
        return fail(s, NBODY_ERR_INVALID_ARG, "radix_bits must be 8, 10 or 11");
    s->radix_bits = t->radix_bits;
    s->walk_block = t->walk_block;
    s->lane_bodies = t->lane_bodies;
    s->wave_bodies = t->wave_bodies;
    s->walk_stack = t->walk_stack;
    s->walk_variant = t->walk_variant;
    s->xcd_order = t->xcd_order;
    s->graph_dirty = true;
    return NBODY_OK;
}

int nbody_get_tuning(const nbody_sim *s, nbody_tuning *out) {
    if (!s || !out) return fail(nullptr, NBODY_ERR_INVALID_ARG, "NULL argument");
    out->walk_block = s->walk_block;
    out->walk_stack = s->walk_stack;
    out->walk_variant = s->walk_variant;
    out->xcd_order = s->xcd_order;
    out->lane_bodies = s->lane_bodies;
    out->wave_bodies = s->wave_bodies;
    out->radix_bits = s->radix_bits;
    return NBODY_OK;
}

int nbody_adjoint_forward(nbody_sim *s, int32_t nsteps, int32_t every) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (nsteps < 1 || every < 1) return fail(s, NBODY_ERR_INVALID_ARG, "nsteps and every must be >= 1");
    if (s->ensemble) return fail(s, NBODY_ERR_INVALID_ARG, "the adjoint supports plain handles only");
    if (s->cfg.softening_mode != NBODY_SOFTEN_CONSTANT)
        return fail(s, NBODY_ERR_INVALID_ARG, "the adjoint needs constant softening");
    if (s->n < 2) return fail(s, NBODY_ERR_INVALID_ARG, "the adjoint needs at least two bodies");
    if (s->cfg.expansion != NBODY_EXPANSION_MONOPOLE)
        return fail(s, NBODY_ERR_INVALID_ARG, "the adjoint differentiates the monopole walk");
    HIP_TRY(s, hipSetDevice(s->device));
    const int64_t n = s->n;
    const int32_t cps = (nsteps + every - 1) / every;
    // (re)allocate: checkpoints, segment positions, cotangents
    if (s->adj_base) {
        HIP_TRY(s, hipStreamSynchronize(s->stream));
        HIP_TRY(s, hipFree(s->adj_base));
        s->adj_base = nullptr;
    }
    Arena a;
    for (int pass = 0; pass < 2; ++pass) {
        a.used = 0;
        s->adj_cp = a.take<float2>((int64_t)cps * 2 * n);
        s->adj_seg = a.take<float2>((int64_t)every * n);
        s->adj_lx = a.take<float2>(n);
        s->adj_lv = a.take<float2>(n);
        s->adj_la = a.take<float2>(n);
        s->adj_end = a.take<float2>(2 * n);
        s->adj_lm = a.take<float>(n);
        s->adj_cot = a.take<float4>(2 * n);
        if (pass == 0) {
            void *p = nullptr;
            hipError_t e = hipMalloc(&p, a.used);
            if (e != hipSuccess) return hip_fail(s, e, "hipMalloc (adjoint)");
            a.base = static_cast<char *>(p);
            s->adj_base = p;
        }
    }
    s->steps_taken += nsteps;
    s->adj_n = n;
    s->adj_steps = nsteps;
    s->adj_every = every;
    s->adj_cps = cps;
    for (int32_t k = 0; k < nsteps; ++k) {
        if (k % every == 0) {
            float2 *cp = s->adj_cp + (size_t)(k / every) * 2 * n;
            HIP_TRY(s, hipMemcpyAsync(cp, s->pos, 8 * (size_t)n, hipMemcpyDeviceToDevice, s->stream));
            HIP_TRY(s, hipMemcpyAsync(cp + n, s->vel, 8 * (size_t)n, hipMemcpyDeviceToDevice, s->stream));
        }
        int rc = enqueue_step(s);
        if (rc != NBODY_OK) return rc;
    }
    return NBODY_OK;
}

int nbody_adjoint_backward(nbody_sim *s, const float *gpos_T, const float *gvel_T, float *gpos0,
                           float *gvel0, float *gmass) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (!s->adj_base || s->adj_n != s->n)
        return fail(s, NBODY_ERR_STATE, "call nbody_adjoint_forward first (same bodies)");
    HIP_TRY(s, hipSetDevice(s->device));
    const int64_t n = s->n;
    const float dt = s->cfg.delta_time;
    hipStream_t st = s->stream;
    // Final state, restored at the end.
    HIP_TRY(s, hipMemcpyAsync(s->adj_end, s->pos, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
    HIP_TRY(s, hipMemcpyAsync(s->adj_end + n, s->vel, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
    if (gpos_T)
        HIP_TRY(s, hipMemcpyAsync(s->adj_lx, gpos_T, 8 * (size_t)n, hipMemcpyHostToDevice, st));
    else
        HIP_TRY(s, hipMemsetAsync(s->adj_lx, 0, 8 * (size_t)n, st));
    if (gvel_T)
        HIP_TRY(s, hipMemcpyAsync(s->adj_lv, gvel_T, 8 * (size_t)n, hipMemcpyHostToDevice, st));
    else
        HIP_TRY(s, hipMemsetAsync(s->adj_lv, 0, 8 * (size_t)n, st));
    HIP_TRY(s, hipMemsetAsync(s->adj_lm, 0, 4 * (size_t)n, st));

    AdjointArgs a{};
    a.leaf = s->tree.leaf;
    a.pcm = s->tree.pcm;
    a.root = s->tree.root;
    a.parent = s->tree.parent;
    a.child = s->tree.child;
    a.la = s->adj_la;
    a.lx = s->adj_lx;
    a.lm = s->adj_lm;
    a.cot = s->adj_cot;
    a.overflow = s->overflow;
    a.n = (uint32_t)n;
    a.grav = s->cfg.grav_constant;
    a.theta = s->cfg.bh_theta;
    open_params(a.theta, a.open_scale, a.open_bias);
    a.eps = s->cfg.epsilon;
    const bool exact = s->cfg.math_mode == NBODY_MATH_EXACT;

    for (int32_t seg = s->adj_cps - 1; seg >= 0; --seg) {
        const int32_t k0 = seg * s->adj_every;
        const int32_t len = std::min(s->adj_every, s->adj_steps - k0);
        // Recompute the segment forward from its checkpoint, keeping x_k.
        const float2 *cp = s->adj_cp + (size_t)seg * 2 * n;
        HIP_TRY(s, hipMemcpyAsync(s->pos, cp, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
        s->bbox_fresh = false;
        HIP_TRY(s, hipMemcpyAsync(s->vel, cp + n, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
        for (int32_t i = 0; i < len; ++i) {
            HIP_TRY(s, hipMemcpyAsync(s->adj_seg + (size_t)i * n, s->pos, 8 * (size_t)n,
                                      hipMemcpyDeviceToDevice, st));
            if (i + 1 < len) {
                int rc = enqueue_step(s);
                if (rc != NBODY_OK) return rc;
            }
        }
        // Reverse sweep over the segment.
        for (int32_t i = len - 1; i >= 0; --i) {
            HIP_TRY(s, hipMemcpyAsync(s->pos, s->adj_seg + (size_t)i * n, 8 * (size_t)n,
                                      hipMemcpyDeviceToDevice, st));
            s->bbox_fresh = false;
            int rc = enqueue_step(s, /*tree_only=*/true);
            if (rc != NBODY_OK) return rc;
            a.idx = s->vals[s->result_buf];
            launch_adjoint_pre(s->adj_lx, s->adj_lv, s->adj_la, (uint32_t)n, dt, st);
            launch_clear_u32(reinterpret_cast<uint32_t *>(s->adj_cot), 4 * (2 * n - 1), st);
            launch_adjoint_walk(a, exact, s->walk_stack, s->lane_bodies, st);
            launch_adjoint_fold(a, st);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(s, e, "adjoint kernel launch");
    if (gpos0) HIP_TRY(s, hipMemcpyAsync(gpos0, s->adj_lx, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
    if (gvel0) HIP_TRY(s, hipMemcpyAsync(gvel0, s->adj_lv, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
    if (gmass) HIP_TRY(s, hipMemcpyAsync(gmass, s->adj_lm, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIP_TRY(s, hipMemcpyAsync(s->pos, s->adj_end, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
    s->bbox_fresh = false;
    HIP_TRY(s, hipMemcpyAsync(s->vel, s->adj_end + n, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
    s->graph_dirty = true;
    return nbody_synchronize(s);
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free step
// nbody_step_static keeps every intermediate buffer of a step in `scratch`,
// a file-scope device array of kStaticScratchBytes that each GPU holds from
// the library's load to the end of the process.  Its address on a device
// comes from hipGetSymbolAddress, once per device (the array's name is a
// host-side stand-in, not a device address); every call on a device shares
// that device's array and clears the words that must start at zero.
namespace {
constexpr size_t kStaticScratchBytes = size_t(16) << 20;
constexpr int kStaticMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {

__global__ void static_extent_kernel(int64_t *sim_off, int64_t n) {
    sim_off[0] = 0;
    sim_off[1] = n;
}

struct StaticDevice {
    std::mutex lock;
    char *base = nullptr;  // this device's `scratch`
    nbody_sim sim;         // buffer views into base; never owns memory
};
StaticDevice g_static[kStaticMaxDevices];

// First call on device `dev` (current): the array's device address, the
// engine defaults, and a check that the largest layout fits the array.
int static_setup(StaticDevice &d, int dev) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetSymbolAddress(scratch)");
    nbody_sim *s = &d.sim;
    s->device = dev;
    s->ensemble = false;
    s->num_sims = 1;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
        s->resident_waves = (uint32_t)cus * 32u;
    s->max_chunks = (int)((NBODY_STATIC_MAX_BODIES + kBBoxChunk - 1) / kBBoxChunk);
    Arena a;
    layout(s, a, NBODY_STATIC_MAX_BODIES, true);
    if (a.used > kStaticScratchBytes)
        return fail(nullptr, NBODY_ERR_CAPACITY,
                    "nbody_step_static: the layout of the largest step needs " +
                        std::to_string(a.used) + " bytes, more than the static scratch");
    d.base = static_cast<char *>(ptr);
    return NBODY_OK;
}

}  // namespace

extern "C" {

int nbody_step_static(const nbody_config *cfg, int64_t n, const float *mass, float *pos_xy,
                      float *vel_xy, int32_t nsteps, void *hip_stream) {
    int rc = validate_config(cfg);
    if (rc != NBODY_OK) return rc;
    if (n < 0 || nsteps < 0)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "nbody_step_static: n and nsteps must be >= 0");
    if (n > NBODY_STATIC_MAX_BODIES)
        return fail(nullptr, NBODY_ERR_CAPACITY,
                    "nbody_step_static: n = " + std::to_string(n) +
                        " exceeds the static scratch limit of " +
                        std::to_string(NBODY_STATIC_MAX_BODIES) + " bodies");
    if (n > 0 && (!mass || !pos_xy || !vel_xy))
        return fail(nullptr, NBODY_ERR_INVALID_ARG,
                    "nbody_step_static: mass/pos/vel must be device pointers");
    if (cfg->expansion != NBODY_EXPANSION_MONOPOLE)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "nbody_step_static: monopole expansion only");
    if (n == 0 || nsteps == 0) return NBODY_OK;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
    if (dev < 0 || dev >= kStaticMaxDevices)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "nbody_step_static: device index too large");
    StaticDevice &d = g_static[dev];
    std::lock_guard<std::mutex> guard(d.lock);
    if (!d.base) {
        rc = static_setup(d, dev);
        if (rc != NBODY_OK) return rc;
    }
    nbody_sim *s = &d.sim;
    s->cfg = *cfg;
    s->key_bits = cfg->morton_mode == NBODY_MORTON_BBOX30 ? 30 : 32;
    s->n = n;
    s->cap = n;
    s->max_chunks = (int)std::max<int64_t>(1, (n + kBBoxChunk - 1) / kBBoxChunk);
    Arena a;
    a.base = d.base;
    layout(s, a, n, true);  // within the largest layout (every size grows with n)
    s->mass = const_cast<float *>(mass);
    s->pos = reinterpret_cast<float2 *>(pos_xy);
    s->vel = reinterpret_cast<float2 *>(vel_xy);
    sort_bind_scratch(s->sort, s->sort_words, n);
    s->stream = static_cast<hipStream_t>(hip_stream);
    s->mass_absmax = NAN;  // masses are on the device: keep the own-leaf compare
    s->bbox_fresh = false;
    s->use_graph = false;
    s->profiling = false;
    s->debug = 0;
    // words that must start at zero (another call may have left them set)
    launch_clear_u32(s->sort_words, sort_scratch_words(n), s->stream);
    launch_clear_u32(s->tree.counts, fill_count_words(n), s->stream);
    launch_clear_u32(reinterpret_cast<uint32_t *>(s->stats), 8, s->stream);
    launch_clear_u32(s->overflow, 1, s->stream);
    static_extent_kernel<<<1, 1, 0, s->stream>>>(s->sim_off, n);
    for (int32_t k = 0; k < nsteps; ++k) {
        rc = enqueue_step(s);
        if (rc != NBODY_OK) {
            g_last_error = s->last_error;
            return rc;
        }
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(nullptr, e, "nbody_step_static: kernel launch");
    return NBODY_OK;
}

int nbody_direct_accel_f64(nbody_sim *s, const int64_t *targets, int64_t ntargets,
                           double *acc_xy) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (s->ensemble) return fail(s, NBODY_ERR_INVALID_ARG, "the direct sum takes plain handles");
    const int64_t n = s->n;
    if (!targets) ntargets = n;
    if (ntargets < 0 || (ntargets > 0 && !acc_xy))
        return fail(s, NBODY_ERR_INVALID_ARG, "bad targets / output");
    std::vector<int64_t> tg((size_t)ntargets);
    for (int64_t k = 0; k < ntargets; ++k) {
        tg[(size_t)k] = targets ? targets[k] : k;
        if (tg[(size_t)k] < 0 || tg[(size_t)k] >= n)
            return fail(s, NBODY_ERR_INVALID_ARG, "target index out of range");
    }
    if (ntargets == 0) return NBODY_OK;
    HIP_TRY(s, hipSetDevice(s->device));
    const int chunks = direct_chunks(n, ntargets);
    const size_t bytes = 8 * (size_t)ntargets + 16 * (size_t)ntargets * ((size_t)chunks + 1);
    char *buf = nullptr;
    HIP_TRY(s, hipMalloc(reinterpret_cast<void **>(&buf), bytes));
    int64_t *d_tg = reinterpret_cast<int64_t *>(buf);
    double2 *d_out = reinterpret_cast<double2 *>(buf + 8 * (size_t)ntargets);
    double2 *d_part = d_out + ntargets;
    hipError_t e = hipMemcpyAsync(d_tg, tg.data(), 8 * (size_t)ntargets, hipMemcpyHostToDevice,
                                  s->stream);
    if (e == hipSuccess) {
        launch_direct_f64(s->pos, s->mass, n, d_tg, ntargets, step_params(s->cfg), d_part, chunks,
                          d_out, s->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(acc_xy, d_out, 16 * (size_t)ntargets, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(s, e, "nbody_direct_accel_f64");
    return NBODY_OK;
}

int nbody_set_radius_multiplier(nbody_sim *s, const float *radius_multiplier) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (s->n > 0 && !radius_multiplier) return fail(s, NBODY_ERR_INVALID_ARG, "NULL array");
    HIP_TRY(s, hipSetDevice(s->device));
    if (s->n > 0)
        HIP_TRY(s, hipMemcpyAsync(s->radius, radius_multiplier, 4 * (size_t)s->n,
                                  hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

int nbody_get_radius_multiplier(nbody_sim *s, float *radius_multiplier) {
    if (!s) return fail(nullptr, NBODY_ERR_INVALID_ARG, "sim is NULL");
    if (s->n > 0 && !radius_multiplier) return fail(s, NBODY_ERR_INVALID_ARG, "NULL array");
    HIP_TRY(s, hipSetDevice(s->device));
    if (s->n > 0)
        HIP_TRY(s, hipMemcpyAsync(radius_multiplier, s->radius, 4 * (size_t)s->n,
                                  hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

int nbody_get_step_count(const nbody_sim *s, int64_t *steps) {
    if (!s || !steps) return fail(nullptr, NBODY_ERR_INVALID_ARG, "NULL argument");
    *steps = s->steps_taken;
    return NBODY_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- snapshots
// A snapshot is a directory of .npy arrays (NumPy format 1.0, little endian,
// C order) plus meta.json:
//   mass.npy f4[n], pos.npy f4[n,2], vel.npy f4[n,2], radius_multiplier.npy
//   f4[n], perm.npy u4[n] (sorted position -> body of the last step; the
//   identity before any step), meta.json {format, n, num_sims, n_per_sim,
//   steps, config}.
namespace {

bool npy_write(const std::string &path, const char *descr, const std::vector<int64_t> &shape,
               const void *data, size_t bytes) {
    std::string dims;
    for (size_t i = 0; i < shape.size(); ++i) dims += std::to_string(shape[i]) + ", ";
    if (shape.size() > 1) dims.erase(dims.size() - 2);
    std::string h = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" +
                    dims + "), }";
    const size_t total = (10 + h.size() + 1 + 63) / 64 * 64;
    h.append(total - 10 - h.size() - 1, ' ');
    h.push_back('\n');
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    const unsigned char pre[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write(reinterpret_cast<const char *>(pre), 8);
    const uint16_t len = (uint16_t)h.size();
    const unsigned char l[2] = {(unsigned char)(len & 0xFF), (unsigned char)(len >> 8)};
    f.write(reinterpret_cast<const char *>(l), 2);
    f.write(h.data(), (std::streamsize)h.size());
    f.write(static_cast<const char *>(data), (std::streamsize)bytes);
    return (bool)f;
}

// Reads a C-order array with the given dtype; false on any mismatch.
bool npy_read(const std::string &path, const char *descr, std::vector<int64_t> &shape,
              std::vector<char> &data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    unsigned char pre[8];
    if (!f.read(reinterpret_cast<char *>(pre), 8) || pre[0] != 0x93 ||
        std::memcmp(pre + 1, "NUMPY", 5) != 0)
        return false;
    uint32_t len = 0;
    if (pre[6] == 1) {
        unsigned char l[2];
        if (!f.read(reinterpret_cast<char *>(l), 2)) return false;
        len = l[0] | (l[1] << 8);
    } else if (pre[6] == 2 || pre[6] == 3) {
        unsigned char l[4];
        if (!f.read(reinterpret_cast<char *>(l), 4)) return false;
        len = l[0] | (l[1] << 8) | (l[2] << 16) | ((uint32_t)l[3] << 24);
    } else {
        return false;
    }
    std::string h(len, '\0');
    if (!f.read(&h[0], len)) return false;
    const std::string want = std::string("'descr': '") + descr + "'";
    if (h.find(want) == std::string::npos || h.find("'fortran_order': False") == std::string::npos)
        return false;
    const size_t a = h.find("'shape': (");
    if (a == std::string::npos) return false;
    const size_t b = h.find(')', a);
    std::stringstream dims(h.substr(a + 10, b - a - 10));
    shape.clear();
    std::string tok;
    size_t count = 1;
    while (std::getline(dims, tok, ',')) {
        if (tok.find_first_not_of(' ') == std::string::npos) continue;
        shape.push_back(std::stoll(tok));
        count *= (size_t)shape.back();
    }
    data.resize(count * 4);
    return (bool)f.read(data.data(), (std::streamsize)data.size());
}

// Number after "key": in a JSON text written by nbody_save_snapshot.
bool json_num(const std::string &t, const std::string &key, double &v) {
    const size_t k = t.find("\"" + key + "\":");
    if (k == std::string::npos) return false;
    v = std::strtod(t.c_str() + k + key.size() + 3, nullptr);
    return true;
}

std::string fmt9(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.9g", v);
    return b;
}

}  // namespace

extern "C" {

int nbody_save_snapshot(nbody_sim *s, const char *dir) {
    if (!s || !dir) return fail(s, NBODY_ERR_INVALID_ARG, "NULL argument");
    const std::string d(dir);
    if (mkdir(dir, 0755) != 0 && errno != EEXIST)
        return fail(s, NBODY_ERR_INVALID_ARG, "cannot create snapshot directory " + d);
    const int64_t n = s->n;
    std::vector<float> mass((size_t)n), pos(2 * (size_t)n), vel(2 * (size_t)n), rad((size_t)n);
    std::vector<uint32_t> perm((size_t)n);
    int rc = nbody_get_bodies(s, mass.data(), pos.data(), vel.data());
    if (rc == NBODY_OK) rc = nbody_get_radius_multiplier(s, rad.data());
    if (rc != NBODY_OK) return rc;
    if (n > 0 && s->steps_taken > 0) {
        HIP_TRY(s, hipMemcpy(perm.data(), s->vals[s->result_buf], 4 * (size_t)n,
                             hipMemcpyDeviceToHost));
    } else {
        for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (uint32_t)i;
    }
    const bool ok = npy_write(d + "/mass.npy", "<f4", {n}, mass.data(), 4 * mass.size()) &&
                    npy_write(d + "/pos.npy", "<f4", {n, 2}, pos.data(), 4 * pos.size()) &&
                    npy_write(d + "/vel.npy", "<f4", {n, 2}, vel.data(), 4 * vel.size()) &&
                    npy_write(d + "/radius_multiplier.npy", "<f4", {n}, rad.data(), 4 * rad.size()) &&
                    npy_write(d + "/perm.npy", "<u4", {n}, perm.data(), 4 * perm.size());
    if (!ok) return fail(s, NBODY_ERR_INVALID_ARG, "cannot write snapshot arrays in " + d);
    const nbody_config &c = s->cfg;
    std::string sizes;
    for (int i = 0; i < s->num_sims; ++i)
        sizes += (i ? ", " : "") + std::to_string(s->sim_off_host[i + 1] - s->sim_off_host[i]);
    std::ofstream m(d + "/meta.json");
    m << "{\"format\": \"nbody_sim snapshot 1\", \"n\": " << n << ", \"num_sims\": " << s->num_sims
      << ", \"n_per_sim\": [" << sizes << "], \"steps\": " << s->steps_taken
      << ", \"config\": {\"grav_constant\": " << fmt9(c.grav_constant)
      << ", \"bh_theta\": " << fmt9(c.bh_theta) << ", \"substeps\": " << c.substeps
      << ", \"delta_time\": " << fmt9(c.delta_time)
      << ", \"epsilon_multiplier\": " << fmt9(c.epsilon_multiplier)
      << ", \"epsilon\": " << fmt9(c.epsilon) << ", \"softening_mode\": " << c.softening_mode
      << ", \"morton_mode\": " << c.morton_mode << ", \"math_mode\": " << c.math_mode
      << ", \"offset\": " << c.offset << "}}\n";
    if (!m) return fail(s, NBODY_ERR_INVALID_ARG, "cannot write " + d + "/meta.json");
    return NBODY_OK;
}

int nbody_load_snapshot(nbody_sim *s, const char *dir) {
    if (!s || !dir) return fail(s, NBODY_ERR_INVALID_ARG, "NULL argument");
    const std::string d(dir);
    std::ifstream mf(d + "/meta.json");
    if (!mf) return fail(s, NBODY_ERR_INVALID_ARG, "no meta.json in " + d);
    std::stringstream ss;
    ss << mf.rdbuf();
    const std::string t = ss.str();
    if (t.find("\"nbody_sim snapshot 1\"") == std::string::npos)
        return fail(s, NBODY_ERR_INVALID_ARG, d + "/meta.json is not an nbody_sim snapshot");
    double n_d = -1, sims_d = 0, steps_d = 0;
    if (!json_num(t, "n", n_d) || !json_num(t, "num_sims", sims_d) || !json_num(t, "steps", steps_d))
        return fail(s, NBODY_ERR_INVALID_ARG, "incomplete meta.json in " + d);
    const int64_t n = (int64_t)n_d;
    if ((int)sims_d != s->num_sims || (s->ensemble && n != s->sim_off_host[s->num_sims]))
        return fail(s, NBODY_ERR_INVALID_ARG, "snapshot layout (sims, bodies) differs from the handle");
    if (s->ensemble) {  // every simulation's size must match
        const size_t a = t.find("\"n_per_sim\": [");
        std::stringstream arr(t.substr(a + 14, t.find(']', a) - a - 14));
        std::string tok;
        for (int i = 0; i < s->num_sims; ++i) {
            if (!std::getline(arr, tok, ',') ||
                std::stoll(tok) != s->sim_off_host[i + 1] - s->sim_off_host[i])
                return fail(s, NBODY_ERR_INVALID_ARG, "snapshot simulation sizes differ");
        }
    }
    nbody_config c = s->cfg;
    double v;
    const std::string cfg = t.substr(t.find("\"config\""));
    if (json_num(cfg, "grav_constant", v)) c.grav_constant = (float)v;
    if (json_num(cfg, "bh_theta", v)) c.bh_theta = (float)v;
    if (json_num(cfg, "substeps", v)) c.substeps = (int32_t)v;
    if (json_num(cfg, "delta_time", v)) c.delta_time = (float)v;
    if (json_num(cfg, "epsilon_multiplier", v)) c.epsilon_multiplier = (float)v;
    if (json_num(cfg, "epsilon", v)) c.epsilon = (float)v;
    if (json_num(cfg, "softening_mode", v)) c.softening_mode = (int32_t)v;
    if (json_num(cfg, "morton_mode", v)) c.morton_mode = (int32_t)v;
    if (json_num(cfg, "math_mode", v)) c.math_mode = (int32_t)v;
    if (json_num(cfg, "offset", v)) c.offset = (uint32_t)v;
    std::vector<int64_t> sh;
    std::vector<char> mass, pos, vel, rad, perm;
    auto need = [&](const char *name, const char *descr, std::vector<char> &buf, int64_t cols) {
        return npy_read(d + "/" + name, descr, sh, buf) && !sh.empty() && sh[0] == n &&
               (cols == 1 ? sh.size() == 1 : (sh.size() == 2 && sh[1] == cols));
    };
    if (!need("mass.npy", "<f4", mass, 1) || !need("pos.npy", "<f4", pos, 2) ||
        !need("vel.npy", "<f4", vel, 2) || !need("radius_multiplier.npy", "<f4", rad, 1) ||
        !need("perm.npy", "<u4", perm, 1))
        return fail(s, NBODY_ERR_INVALID_ARG, "missing or malformed arrays in " + d);
    int rc = nbody_set_config(s, &c);
    if (rc == NBODY_OK)
        rc = nbody_set_bodies(s, n, reinterpret_cast<const float *>(mass.data()),
                              reinterpret_cast<const float *>(pos.data()),
                              reinterpret_cast<const float *>(vel.data()));
    if (rc == NBODY_OK) rc = nbody_set_radius_multiplier(s, reinterpret_cast<const float *>(rad.data()));
    if (rc != NBODY_OK) return rc;
    if (n > 0) {  // the last step's permutation, for nbody_debug_get
        HIP_TRY(s, hipMemcpy(s->vals[0], perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
        s->result_buf = 0;
    }
    s->steps_taken = (int64_t)steps_d;
    return NBODY_OK;
}

int nbody_default_camera(int32_t width, int32_t height, nbody_camera *out) {
    if (!out || width < 1 || height < 1)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "camera needs width, height >= 1");
    // renderer.ts:49-50, 240-243: centre (0, 0), half height 10, half width
    // 10 * aspect
    out->center_x = 0.0f;
    out->center_y = 0.0f;
    out->half_y = 10.0f;
    out->half_x = 10.0f * ((float)width / (float)height);
    out->width = width;
    out->height = height;
    return NBODY_OK;
}

int nbody_render(nbody_sim *s, const nbody_camera *cam, uint16_t *density_f16, uint8_t *rgba) {
    if (!s || !cam) return fail(s, NBODY_ERR_INVALID_ARG, "NULL argument");
    if (cam->width < 1 || cam->height < 1 || cam->width > 16384 || cam->height > 16384)
        return fail(s, NBODY_ERR_INVALID_ARG, "viewport must be 1..16384 pixels per side");
    if (!(cam->half_x > 0.0f) || !(cam->half_y > 0.0f))
        return fail(s, NBODY_ERR_INVALID_ARG, "camera half sizes must be > 0");
    HIP_TRY(s, hipSetDevice(s->device));
    const int64_t px = (int64_t)cam->width * cam->height;
    if (cam->width != s->render_w || cam->height != s->render_h) {
        HIP_TRY(s, hipStreamSynchronize(s->stream));
        if (s->render_base) HIP_TRY(s, hipFree(s->render_base));
        s->render_base = nullptr;
        Arena a;
        for (int pass = 0; pass < 2; ++pass) {
            a.used = 0;
            s->render_dens = a.take<unsigned long long>(px);
            s->render_d16 = a.take<uint16_t>(px);
            s->render_rgba = a.take<uint32_t>(px);
            s->render_lut = a.take<uint32_t>(65536);
            if (pass == 0) {
                void *p = nullptr;
                HIP_TRY(s, hipMalloc(&p, a.used));
                a.base = static_cast<char *>(p);
                s->render_base = p;
            }
        }
        s->render_w = cam->width;
        s->render_h = cam->height;
        s->render_lut_ready = false;
    }
    RenderParams rp;
    rp.cam_x = cam->center_x;
    rp.cam_y = cam->center_y;
    rp.half_x = cam->half_x;
    rp.half_y = cam->half_y;
    rp.width = cam->width;
    rp.height = cam->height;
    const float zoom = std::max(cam->half_x, cam->half_y);  // density.wgsl:86-88
    rp.zoom_scale = 1.0f / (1.0f + 0.05f * zoom);
    launch_render(s->pos, s->radius, (uint32_t)s->n, rp, s->render_dens, s->render_lut,
                  !s->render_lut_ready, s->render_d16, s->render_rgba, s->stream);
    HIP_TRY(s, hipGetLastError());
    s->render_lut_ready = true;
    if (density_f16)
        HIP_TRY(s, hipMemcpyAsync(density_f16, s->render_d16, 2 * (size_t)px,
                                  hipMemcpyDeviceToHost, s->stream));
    if (rgba)
        HIP_TRY(s, hipMemcpyAsync(rgba, s->render_rgba, 4 * (size_t)px, hipMemcpyDeviceToHost,
                                  s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

int nbody_write_png(const char *path, int32_t width, int32_t height, const uint8_t *rgba) {
    if (!path || !rgba || width < 1 || height < 1 || width > 65535 || height > 65535)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "nbody_write_png: bad arguments");
    // scanlines with filter type 0 (none) in front of each row
    const size_t row = 4 * (size_t)width;
    std::vector<unsigned char> raw((row + 1) * (size_t)height);
    for (int32_t y = 0; y < height; ++y) {
        raw[(row + 1) * (size_t)y] = 0;
        std::memcpy(&raw[(row + 1) * (size_t)y + 1], rgba + row * (size_t)y, row);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return fail(nullptr, NBODY_ERR_INVALID_ARG, "nbody_write_png: zlib failed");
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(nullptr, NBODY_ERR_INVALID_ARG, std::string("cannot write ") + path);
    auto be32 = [](unsigned char *p, uint32_t v) {
        p[0] = (unsigned char)(v >> 24);
        p[1] = (unsigned char)(v >> 16);
        p[2] = (unsigned char)(v >> 8);
        p[3] = (unsigned char)v;
    };
    auto chunk = [&](const char *type, const unsigned char *data, uint32_t len) {
        unsigned char hdr[8];
        be32(hdr, len);
        std::memcpy(hdr + 4, type, 4);
        f.write(reinterpret_cast<const char *>(hdr), 8);
        if (len) f.write(reinterpret_cast<const char *>(data), len);
        uLong crc = crc32(0L, reinterpret_cast<const Bytef *>(type), 4);
        if (len) crc = crc32(crc, data, len);
        unsigned char c[4];
        be32(c, (uint32_t)crc);
        f.write(reinterpret_cast<const char *>(c), 4);
    };
    const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    f.write(reinterpret_cast<const char *>(sig), 8);
    unsigned char ihdr[13];
    be32(ihdr, (uint32_t)width);
    be32(ihdr + 4, (uint32_t)height);
    ihdr[8] = 8;   // bits per channel
    ihdr[9] = 6;   // RGBA
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering (each row says none)
    ihdr[12] = 0;  // no interlace
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), (uint32_t)zlen);
    chunk("IEND", nullptr, 0);
    if (!f) return fail(nullptr, NBODY_ERR_INVALID_ARG, std::string("cannot write ") + path);
    return NBODY_OK;
}

const char *nbody_last_error(const nbody_sim *s) {
    return s ? s->last_error.c_str() : g_last_error.c_str();
}

const char *nbody_version(void) { return "nbody_sim 0.1.0 (gfx950)"; }

}  // extern "C"
