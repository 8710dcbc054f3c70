// nbody_internal.h -- device-side types and kernel launchers shared by the
// HIP translation units of libnbody_sim.so.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbs {

constexpr int kWave = 64;
// Waits of one workgroup on another (the sort's look-back, the fill's done
// flags) are bounded: the workgroups they wait for are always dispatched
// first and wait on nothing, so a wait ends in microseconds; one that polls
// kSpinLimit times (seconds) stops, and sets kFaultWait in the handle's
// fault word, which nbody_synchronize reports, rather than hang the GPU.
constexpr uint32_t kSpinLimit = 1u << 24;
constexpr uint32_t kFaultWait = 2u;

// Lanes of mask m below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Exclusive scan of one value per thread over a kThreads-thread block
// (every thread calls it; shw holds kThreads / 64 words).
template <int kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *shw) {
    constexpr int kWaves = kThreads / kWave;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) shw[wid] = x;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < kWaves; ++w)
        if (w < wid) before += shw[w];
    __syncthreads();
    return before + x - v;
}
constexpr uint32_t kNoNode = 0xFFFFFFFFu;
constexpr int kSimShift = 30;     // ensemble keys: sim id << 30 | 30-bit Morton code
constexpr int kMaxSims = 256;     // 8-bit sim prefix
constexpr int kBBoxChunk = 4096;  // bodies per bounding-box partial
constexpr int kBBoxSlots = 256;    // bounding-box accumulators filled by the walk

// Order-preserving u32 image of a binary32 number (larger number, larger
// image; -0 below +0) for atomic min/max.  Images of numbers are never 0,
// so 0 stands for "no value": the bounding-box accumulators hold the
// maximum of image(x) for the maximum and of ~image(x) for the minimum.
__host__ __device__ inline uint32_t float_order(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float float_unorder(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __builtin_bit_cast(float, b);
}

// Per-step physics parameters (the Metadata uniform, lbvh.wgsl:3-10).
struct StepParams {
    float grav;
    float dt;
    float eps_mult;
    float theta;
    float eps;  // constant softening length
    int soft_mode;
    int morton_mode;
    int math_mode;
    int expansion;  // 0 monopole, 1 quadrupole (second moments in the pair records)
    float open_scale, open_bias;  // the fast arithmetic's opening keys (open_key)
};

// The fast arithmetic's opening test.  A body passes a node of length len
// (the reference's length * inv < theta, barnes_hut.wgsl:89) when
// !(open_key(len) >= d2), d2 = rx^2 + ry^2 + eps^2 the squared softened
// distance: key = len |len| / theta^2, so the test needs no square root and
// the walks take the reciprocal square root only for the nodes some body
// accepts.  In exact arithmetic both tests agree; in binary32 they part only
// where len / d lies within a few units in the last place of theta.  Leaves
// (length -inf) get the key -inf, or NaN when theta <= 0: both pass for every
// d2.  theta <= 0 (or NaN) makes the bias +inf, so every internal node opens,
// as the reference's test does for such a theta.  A body whose d2 is NaN (a
// NaN position) passes every node: its pull is NaN, as in the reference,
// which opens every node for it instead.  The pair records carry the
// children's keys in their first two padding words (monopole), written by the
// fill with this same expression.
__host__ __device__ inline float open_key(float len, float scale, float bias) {
    return len * fabsf(len) * scale + bias;
}
inline void open_params(float theta, float &scale, float &bias) {
    if (theta > 0.0f) {
        double s = 1.0 / ((double)theta * (double)theta);
        s = s < 1.1754943508222875e-38 ? 1.1754943508222875e-38 : s;  // FLT_MIN: never 0
        s = s > 3.4028234663852886e38 ? 3.4028234663852886e38 : s;    // FLT_MAX: never inf
        scale = (float)s;
        bias = 0.0f;
    } else {
        scale = 3.4028234663852886e38f;
        bias = __builtin_huge_valf();
    }
}

// ---------------------------------------------------------------- sort scratch
constexpr int kSortMaxPasses = 5;     // 38-bit ensemble keys at 8 bits per digit
constexpr int kSortHistWords = kSortMaxPasses * 2048;  // [pass][2^radix_bits] digit counts
// The persistent walk's queue counter: a word of the sort's counter block
// (zeroed with the sort's scratch by the build; the sort uses words
// [0, kSortMaxPasses)).
constexpr int kWalkQueueWord = 16;
struct SortScratch {
    uint32_t *hist;      // [passes][2^radix_bits] digit counts (written by the encode)
    uint32_t *lookback;  // [passes][tiles][2^radix_bits] decoupled look-back words
    uint32_t *counters;  // [passes] dynamic tile ids
    int64_t words;       // u32 words of the whole scratch
};
int sort_tile_for(int64_t n, bool wide_keys);  // keys per one-sweep tile (64-bit keys: wide)
int64_t sort_tiles(int64_t n);                 // the most tiles of any tile size
int sort_passes(int key_bits, int radix_bits);
int64_t sort_scratch_words(int64_t cap);  // enough for every radix width and tile size
// Words to clear (from sc.hist on) after a sort of n keys, for the next one.
int64_t sort_clear_words(int64_t n, int key_bits, int radix_bits, bool wide_keys);
void sort_bind_scratch(SortScratch &sc, uint32_t *base, int64_t cap);

// ---------------------------------------------------------------- keys
// bbox partials: grid (chunks, sims); then one wave per sim reduces them.
// slots != nullptr (one sim): also set the walk's accumulators to this box.
void launch_bbox(const float2 *pos, const int64_t *sim_off, int num_sims, int max_chunks,
                 float4 *partial, float4 *bbox, uint32_t *slots, hipStream_t s);
// Morton encode, keys32 (plain) or keys64 (ensemble, body_sim != nullptr),
// counting the digits of every sort pass into hist ([passes][2^radix_bits],
// which must be zero; passes = 0 counts nothing).  slots != nullptr: the
// bounding box comes from the walk's accumulators (and is stored in bbox[0]).
void launch_encode(const float2 *pos, int64_t n, int morton_mode, float4 *bbox,
                   const uint32_t *slots, const uint8_t *body_sim, uint32_t *keys32,
                   uint64_t *keys64, uint32_t *idx, uint32_t *hist, int passes, int radix_bits,
                   hipStream_t s);
// The same keys, grouped by bucket inside each tile of encode_group_tile(n)
// bodies (stable; keys32/idx at the tile's positions), with every tile's
// count and first position of each bucket in tcount / tstart (bucket-major,
// nb x encode_group_tiles(n) words each) and the counts added to hist[0..nb).
void launch_encode_group(const float2 *pos, int64_t n, int morton_mode, float4 *bbox,
                         const uint32_t *slots, uint32_t *keys32, uint32_t *idx,
                         const uint32_t *split, int nb, uint32_t *hist, uint32_t *tcount,
                         uint32_t *tstart, hipStream_t s);
int encode_group_tiles(int64_t n);
int encode_group_tile(int64_t n);

// ---------------------------------------------------------------- sort
// Stable LSD sort of (keys, vals) on the low `key_bits` bits, radix_bits
// (8, 10 or 11) per pass; the digit histograms must be in sc.hist.
// Ping-pongs between buffers a and b and returns which buffer (0 = a, 1 = b)
// holds the result.
// The last pass also writes each body's sorted position, rank[vals[t]] = t,
// when `rank` is set (the build then places the leaf records).
struct LeafGather {
    uint32_t *rank;
    uint32_t *fault = nullptr;         // kFaultWait is or-ed in when a look-back wait gives up
};
int launch_sort_u32(uint32_t *keys_a, uint32_t *vals_a, uint32_t *keys_b, uint32_t *vals_b,
                    int64_t n, int key_bits, int radix_bits, const SortScratch &sc,
                    const LeafGather &lg, hipStream_t s);
// The bucket sort of plain handles (bit-identical to the radix sort): one
// workgroup sorts n <= kSortLdsMax keys (launch_sort_small, result in the b
// buffers); larger sorts up to kMaxBuckets << kBucketShiftMax keys split them
// between the previous step's splitters (launch_encode_group groups each
// tile's keys by bucket, launch_sort_gathered sorts every bucket).
constexpr int64_t kSortLdsMax = 4096;
constexpr int kBucketShiftMax = 12;  // widest splitter spacing: 4096 sorted positions
constexpr int kMaxBuckets = 1024;
constexpr int kBucketCap = 6144;     // largest bucket sorted in LDS (larger: through memory)
// The splitter spacing for n keys: 2^k less 1/32 of it for k = 10 or 11,
// the narrowest that makes at most 256 buckets (small sorts want many
// workgroups), else 4096 less 1/32 while that makes at most kMaxBuckets,
// else 4096.  A bucket holds about that many keys, give or take the few
// bodies that crossed a splitter since the last step, and the 1/32 keeps it
// within 2^k, the size the per-bucket sort handles in whole rows of 1024
// (one more key would cost a whole row more).  Below 2^18 keys the spacing
// is at most 1984, so at most 256 buckets (the encode's 1024-body tiles rank
// 8 bits of bucket id).
__host__ __device__ inline int bucket_spacing(int64_t n) {
    for (int sh = 10; sh < kBucketShiftMax; ++sh) {
        const int s = (1 << sh) - (1 << sh) / 32;
        if ((n + s - 1) / s <= 256) return s;
    }
    const int full = 1 << kBucketShiftMax, s = full - full / 32;
    return (n + s - 1) / s <= kMaxBuckets ? s : full;
}
__host__ __device__ inline int sort_buckets(int64_t n) {
    const int s = bucket_spacing(n);
    return (int)((n + s - 1) / s);
}
uint32_t sort_key_max(int key_bits);
int launch_sort_small(uint32_t *keys_a, uint32_t *vals_a, uint32_t *keys_b, uint32_t *vals_b,
                      int64_t n, int key_bits, const LeafGather &lg, hipStream_t s);
// The per-bucket sort of launch_encode_group's output: bucket b's keys
// gathered tile after tile from (keys_in, vals_in), sorted to (keys_out,
// vals_out) at the bucket's start (the counts of the buckets before it, from
// hist); buckets too large for LDS are sorted between their output positions
// and the same positions of (keys_spill, vals_spill).  Returns -1 (nothing
// launched) for more than 2^22 keys.
int launch_sort_gathered(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out,
                         uint32_t *vals_out, uint32_t *keys_spill, uint32_t *vals_spill, int64_t n,
                         int key_bits, int nb, const uint32_t *split, const uint32_t *hist,
                         const uint32_t *tcount, const uint32_t *tstart, const LeafGather &lg,
                         hipStream_t s);
int launch_sort_u64(uint64_t *keys_a, uint32_t *vals_a, uint64_t *keys_b, uint32_t *vals_b,
                    int64_t n, int key_bits, int radix_bits, const SortScratch &sc,
                    const LeafGather &lg, hipStream_t s);

// ---------------------------------------------------------------- tree
// Node data is stored in child pairs: each internal node p has a 64-byte
// pair record pcm[4p .. 4p+3]: the {com.x, com.y, mass, length} records of
// its left and right children, then the two child indices {left, right}
// (bits of pcm[4p+2].x, .y; the rest is padding).  The children's AABBs sit
// in paabb[2p], paabb[2p+1].  Opening a node therefore fetches everything
// the walk needs from one aligned 64-byte line.  The root, which is
// nobody's child, has its own record.
constexpr uint32_t kRightChild = 0x80000000u;  // parent word: index | flags | side << 31
// Length field of a leaf's node record.  Leaves are accepted whatever the
// opening test says (barnes_hut.wgsl:89); with -inf as their length, the
// walks' test passes for every leaf, so they need no leaf check.  The
// reference's leaf length is 0 (the host reports it so).
constexpr float kLeafLength = -__builtin_huge_valf();
// Parent word flag: the parent's leaf range lies in one 1024-leaf block
// (the fill completes it in its first phase).  Parent indices are < 2^30.
constexpr uint32_t kInBlock = 0x40000000u;
constexpr uint32_t kParentMask = 0x3FFFFFFFu;
constexpr int kPairStride = 4;  // float4s per pair record
// Slot of a node's record in pcm / in paabb, from its parent word.
__host__ __device__ inline size_t pcm_slot(uint32_t pw) {
    return (size_t)kPairStride * (pw & kParentMask) + (pw >> 31);
}
__host__ __device__ inline size_t paabb_slot(uint32_t pw) {
    return 2 * (size_t)(pw & kParentMask) + (pw >> 31);
}
struct TreeBufs {
    uint2 *child;        // [n-1] {left, right}
    uint2 *range;        // [n-1] {first, last} sorted leaf of each internal node
    uint32_t *parent;    // [2n-1] parent index | kRightChild for right children
    uint32_t *status;    // [n-1] arrival counters of the fill
    float4 *leaf;        // [n] sorted bodies {x, y, mass, 0}
    float4 *pcm;         // [4(n-1)] pair records (children's {com, mass, length}, indices)
    float4 *paabb;       // [2(n-1)] children's {min.x, min.y, max.x, max.y}
    float4 *root;        // [3] root {cm, aabb, second moments {xx, xy, yy, 0}}
    uint32_t *counts;    // fill counters (fill_count_words)
    uint32_t *lists;     // the fill's top node list (fill_list_words)
    uint32_t *slot;      // [n-1] position of a top node in the top list
    uint32_t *fault;     // kFaultWait is or-ed in when a wait on another workgroup gives up
    // with pair_keys set, the fill also writes each child's opening key
    // open_key(length, open_scale, open_bias) into its pair record's padding
    // (word 10 left, 11 right)
    uint32_t pair_keys;
    float open_scale, open_bias;
};
// The builds also write the sorted leaf records t.leaf[rank[b]] = {pos[b],
// mass[b], 0} (body order in, scattered out), zero `clear_words` words at
// `clear` (the sort's scratch, for the next step's encode and sort) and the
// kBBoxSlots x 4 words at `slots` (the walk's bounding-box accumulators; may
// be null).
struct LeafPlace {
    const float2 *pos;
    const float *mass;
    const uint32_t *rank;
    uint32_t *split = nullptr;  // u32 builds: the next sort's splitters (keys at j * split_every)
    int split_every = 1 << kBucketShiftMax;
};
void launch_build_u32(const uint32_t *keys, int64_t n, const TreeBufs &t, const LeafPlace &lp,
                      uint32_t *clear, int64_t clear_words, uint32_t *slots, hipStream_t s);
void launch_build_u64(const uint64_t *keys, int64_t n, const int64_t *sim_off, const TreeBufs &t,
                      const LeafPlace &lp, uint32_t *sim_root, uint32_t *clear,
                      int64_t clear_words, uint32_t *slots, hipStream_t s);
void launch_fill(const uint32_t *idx, const float *mass, const float2 *pos, int64_t n,
                 const TreeBufs &t, bool climb, hipStream_t s);
int64_t fill_count_words(int64_t cap);  // sizes of counts / top list for `cap` bodies
int64_t fill_list_words(int64_t cap);

// ---------------------------------------------------------------- walk
struct WalkArgs {
    const float4 *leaf;      // [n] sorted bodies {x, y, mass, 0}
    const float4 *pcm;       // [4(n-1)] pair records
    const float4 *root;      // root record
    const uint32_t *parent;  // parent words (ensemble sub-tree roots)
    const uint2 *child;      // [n-1]
    const uint32_t *idx;     // [n] sorted position -> body
    const uint64_t *keys64;  // ensemble: sorted keys (sim id in bits 30..37)
    const uint32_t *sim_root;
    float2 *vel;             // [n] by body
    float2 *pos;             // [n] by body (drift output)
    float2 *accel;           // optional [n] by body
    unsigned long long *stats;  // optional {accepted, opened, overflowed}
    uint2 *body_counts;         // optional [n] by body {accepted, opened} (pair walks, stats mode)
    uint32_t *overflow;         // set to 1 when a full stack drops nodes
    uint32_t n;
    // 1 when a body's own leaf adds an exact zero to its fast-mode sum:
    // constant softening with eps^2 and the largest |mass| such that
    // |m| * eps^-3 stays finite (the offset is exactly zero)
    uint32_t self_zero;
    // bounding-box accumulators of the drifted positions (the next step's
    // keys), or null
    uint32_t *bbox_slots;
    // persistent work-queue walk: the queue counter (zero at launch), the
    // chunks' order (null: sorted order) and the wavefronts the GPU holds
    uint32_t *queue;
    const uint32_t *chunk_order;
    uint32_t resident;
    uint32_t wave_bodies;  // pair walks: bodies per wavefront (64, 32, 16, 8)
    StepParams p;
};
void launch_walk(const WalkArgs &a, bool ensemble, bool stats, int block, int stack,
                 int xcd_order, int variant, int lane_bodies, hipStream_t s);

// ---------------------------------------------------------------- adjoint
struct AdjointArgs {
    const float4 *leaf;
    const float4 *pcm;
    const float4 *root;
    const uint32_t *parent;
    const uint2 *child;
    const uint32_t *idx;
    const float2 *la;   // [n] dL/da by body
    float2 *lx;         // [n] dL/dx by body (accumulated)
    float *lm;          // [n] dL/dm by body (accumulated)
    float4 *cot;        // [2n-1] node cotangents {lc.x, lc.y, lM, -}
    uint32_t *overflow;
    uint32_t n;
    float grav, theta, eps;
    float open_scale, open_bias;  // open_key's parameters for theta
};
void launch_adjoint_pre(const float2 *lx, float2 *lv, float2 *la, uint32_t n, float dt,
                        hipStream_t s);
void launch_adjoint_walk(const AdjointArgs &a, bool exact, int stack, int lane_bodies,
                         hipStream_t s);
void launch_adjoint_fold(const AdjointArgs &a, hipStream_t s);

// ---------------------------------------------------------------- FP64 direct sum
// Softened direct sum in binary64 for the given target bodies (the walk's
// accuracy reference): partial[chunks][ntargets], out[ntargets].
int direct_chunks(int64_t n, int64_t ntargets);
void launch_direct_f64(const float2 *pos, const float *mass, int64_t n, const int64_t *targets,
                       int64_t ntargets, const StepParams &p, double2 *partial, int chunks,
                       double2 *out, hipStream_t s);

// ---------------------------------------------------------------- quadrupole
// Second mass moments of every node (k_quad.hip) into the pair records'
// padding and root[2]; `prefix` holds quad_prefix_words(n) doubles.
int64_t quad_prefix_words(int64_t cap);
void launch_quad(const TreeBufs &tb, int64_t n, double *prefix, hipStream_t s);

// ---------------------------------------------------------------- render
// Camera and viewport of the density splat (the Renderer's Metadata,
// density.wgsl:3-10) and the splat's zoom factor 1 / (1 + 0.05 max(half)).
struct RenderParams {
    float cam_x, cam_y, half_x, half_y;
    int width, height;
    float zoom_scale;
};
// Clears dens (u64 fixed point, 2^-40 units, [height][width] texture rows),
// builds the tone table lut[65536] when build_lut, splats n bodies, then
// writes the fp16 density (d16, texture rows; may be null) and the RGBA8
// image (rgba, output rows: row y shows texture row height-1-y).
void launch_render(const float2 *pos, const float *radius, uint32_t n, const RenderParams &rp,
                   unsigned long long *dens, uint32_t *lut, bool build_lut, uint16_t *d16,
                   uint32_t *rgba, hipStream_t s);

// ---------------------------------------------------------------- misc
void launch_clear_u32(uint32_t *p, int64_t words, hipStream_t s);
void launch_fill_f32(float *p, float value, int64_t count, hipStream_t s);

}  // namespace nbs
