// k_tree.hip -- Karras LBVH construction and the bottom-up mass/COM fill.
//
// build: build_lbvh_main (src/shaders/compute/lbvh.wgsl:85-165), Karras 2012
//   Fig. 4, one thread per internal node.  Internal nodes are [0, n-2],
//   leaves [n-1, 2n-2] (lbvh.wgsl:94).  Children go to a packed uint2 array
//   (8 B per internal node, what the walk loads), parents to their own array.
// fill: fill_lbvh_main (lbvh.wgsl:168-233).  One thread per leaf in sorted
//   order climbs toward the root; the first thread to reach a parent stops,
//   the second merges both children.  The reference's arrival counter is a
//   relaxed atomic with workgroup-scope barriers, so the second arriver may
//   read stale sibling data (README.md:18).
//
//   Agent-scope acquire/release removes the race but, on a multi-XCD MI355X,
//   writes back and invalidates the XCD's L2 at every counter update; done by
//   every thread at every level it costs milliseconds.  The climb is split in
//   three phases so that every counter is ordered at workgroup scope:
//     phase 1 (fill_local_kernel): workgroup b owns sorted leaves
//       [b*1024, b*1024+1023].  It loads its slice of the child, range and
//       parent arrays into LDS, completes every node whose leaf range lies in
//       the block with LDS-resident node data and LDS arrival counters, then
//       writes all completed nodes out cooperatively.  A node whose parent
//       spans blocks goes to its super-block's (32 blocks) frontier bucket.
//     phase 2 (fill_mid_kernel): one workgroup per super-block climbs from
//       its bucket through the nodes inside the super-block (global arrival
//       counters, workgroup scope); nodes whose parent spans super-blocks go
//       to the top list.
//     phase 3 (fill_top_kernel): one workgroup finishes the top of the tree.
//   Kernel boundaries order the phases.
//   The merge formula does not depend on which child arrives last, so the
//   result is deterministic and equals a sequential post-order fill.  The
//   reference's random leaf rotation (`offset`, lbvh.wgsl:171) is therefore
//   unnecessary and not applied.
#include "nbody_internal.h"

namespace nbs {

namespace {

struct DeltaU32 {
    const uint32_t *__restrict__ k;
    int64_t n;
    // lbvh.wgsl:72-83
    __device__ __forceinline__ int operator()(int64_t i, int64_t j) const {
        if (j < 0 || j >= n) return -1;
        const uint32_t a = k[i], b = k[j];
        if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
        return __clz((int)(a ^ b));
    }
};

// Ensemble keys carry the sim id above bit 30.  Equal keys share a sim and
// are tie-broken on sim-local sorted positions, so every sim's sub-tree has
// the shape of the tree a plain run of that sim builds.
struct DeltaU64 {
    const uint64_t *__restrict__ k;
    const int64_t *__restrict__ sim_off;
    int64_t n;
    __device__ __forceinline__ int operator()(int64_t i, int64_t j) const {
        if (j < 0 || j >= n) return -1;
        const uint64_t a = k[i], b = k[j];
        if (a == b) {
            const int64_t off = sim_off[a >> kSimShift];
            return 64 + __clz((int)((uint32_t)(i - off) ^ (uint32_t)(j - off)));
        }
        return __clzll((long long)(a ^ b));
    }
};

struct Split {
    int64_t lo, hi, gamma;
    uint32_t lc, rc;
};

template <typename Delta>
__device__ __forceinline__ Split karras_node(const Delta &delta, int64_t i, int64_t n) {
    // lbvh.wgsl:101-152
    const int d = delta(i, i + 1) > delta(i, i - 1) ? 1 : -1;
    const int delta_min = delta(i, i - d);
    int64_t l_max = 2;
    while (delta(i, i + d * l_max) > delta_min) l_max *= 2;
    int64_t l = 0;
    for (int64_t t = l_max / 2; t >= 1; t /= 2)
        if (delta(i, i + d * (l + t)) > delta_min) l += t;
    const int64_t j = i + d * l;
    const int delta_node = delta(i, j);
    int64_t s = 0;
    int64_t t = (l + 1) / 2;
    for (;;) {
        if (delta(i, i + d * (s + t)) > delta_node) s += t;
        if (t == 1) break;
        t = (t + 1) / 2;
    }
    Split r;
    r.gamma = i + d * s + (d < 0 ? d : 0);
    r.lo = i < j ? i : j;
    r.hi = i < j ? j : i;
    r.lc = (r.lo == r.gamma) ? (uint32_t)(n - 1 + r.gamma) : (uint32_t)r.gamma;
    r.rc = (r.hi == r.gamma + 1) ? (uint32_t)(n - 1 + r.gamma + 1) : (uint32_t)(r.gamma + 1);
    return r;
}

__global__ __launch_bounds__(256) void build_u32_kernel(const uint32_t *__restrict__ keys,
                                                        int64_t n, TreeBufs tb) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    DeltaU32 delta{keys, n};
    const Split sp = karras_node(delta, i, n);
    tb.child[i] = make_uint2(sp.lc, sp.rc);
    // the walk reads the indices from the pair record
    *reinterpret_cast<uint2 *>(tb.pcm + kPairStride * (size_t)i + 2) = make_uint2(sp.lc, sp.rc);
    tb.range[i] = make_uint2((uint32_t)sp.lo, (uint32_t)sp.hi);
    tb.parent[sp.lc] = (uint32_t)i;  // lbvh.wgsl:155-161
    tb.parent[sp.rc] = (uint32_t)i | kRightChild;
    if (i == 0) tb.parent[0] = kNoNode;
    tb.status[i] = 0u;  // lbvh.wgsl:164
}

__global__ __launch_bounds__(256) void build_u64_kernel(const uint64_t *__restrict__ keys,
                                                        int64_t n,
                                                        const int64_t *__restrict__ sim_off,
                                                        TreeBufs tb,
                                                        uint32_t *__restrict__ sim_root) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1) return;
    DeltaU64 delta{keys, sim_off, n};
    const Split sp = karras_node(delta, i, n);
    tb.child[i] = make_uint2(sp.lc, sp.rc);
    // the walk reads the indices from the pair record
    *reinterpret_cast<uint2 *>(tb.pcm + kPairStride * (size_t)i + 2) = make_uint2(sp.lc, sp.rc);
    tb.range[i] = make_uint2((uint32_t)sp.lo, (uint32_t)sp.hi);
    tb.parent[sp.lc] = (uint32_t)i;
    tb.parent[sp.rc] = (uint32_t)i | kRightChild;
    tb.status[i] = 0u;
    // A node spanning two sims sits above the sim level; a child whose
    // range lies inside one sim is that sim's root.
    const uint64_t s_lo = keys[sp.lo] >> kSimShift, s_hi = keys[sp.hi] >> kSimShift;
    if (s_lo != s_hi) {
        if ((keys[sp.gamma] >> kSimShift) == s_lo) sim_root[s_lo] = sp.lc;
        if ((keys[sp.gamma + 1] >> kSimShift) == s_hi) sim_root[s_hi] = sp.rc;
    }
    if (i == 0) {
        tb.parent[0] = kNoNode;
        const uint64_t s0 = keys[0] >> kSimShift;
        if (s0 == (keys[n - 1] >> kSimShift)) sim_root[s0] = 0u;
    }
}

constexpr int kFillBlock = 1024;

// lbvh.wgsl:205-228: a node from its two complete children.
__device__ __forceinline__ void merge_values(const float4 &L, const float4 &LB, const float4 &R,
                                             const float4 &RB, float4 &C, float4 &B) {
    B.x = fminf(LB.x, RB.x);
    B.y = fminf(LB.y, RB.y);
    B.z = fmaxf(LB.z, RB.z);
    B.w = fmaxf(LB.w, RB.w);
    C.z = L.z + R.z;
    if (C.z > 0.0f) {
        C.x = (L.x * L.z + R.x * R.z) / C.z;
        C.y = (L.y * L.z + R.y * R.z) / C.z;
    } else {
        C.x = 0.0f;
        C.y = 0.0f;
    }
    C.w = fmaxf(B.z - B.x, B.w - B.y);
}

// Node p from its children's records in p's child-pair slots.
__device__ __forceinline__ void merge_children(const TreeBufs &tb, uint32_t p, float4 &C,
                                               float4 &B) {
    merge_values(tb.pcm[kPairStride * (size_t)p], tb.paabb[2 * (size_t)p],
                 tb.pcm[kPairStride * (size_t)p + 1], tb.paabb[2 * (size_t)p + 1], C, B);
}

// Store a completed node's data where its readers look for it: the child
// slot of its parent (parent word `pw`), or the root record.
__device__ __forceinline__ void publish(const TreeBufs &tb, uint32_t node, uint32_t pw,
                                        const float4 &C, const float4 &B) {
    if (node == 0u) {
        tb.root[0] = C;
        tb.root[1] = B;
        return;
    }
    tb.pcm[pcm_slot(pw)] = C;
    tb.paabb[paabb_slot(pw)] = B;
}

constexpr int kSuperBlocks = 32;                     // phase-1 blocks per phase-2 workgroup
constexpr int kSuper = kFillBlock * kSuperBlocks;   // leaves per super-block

__global__ __launch_bounds__(kFillBlock) void fill_local_kernel(const uint32_t *__restrict__ idx,
                                                                const float *__restrict__ mass,
                                                                const float2 *__restrict__ pos,
                                                                int64_t n, TreeBufs tb) {
    __shared__ uint2 s_child[kFillBlock];     // internal node base+j
    __shared__ uint2 s_range[kFillBlock];
    __shared__ uint32_t s_par_int[kFillBlock];
    __shared__ uint32_t s_par_leaf[kFillBlock];  // leaf of sorted position base+j
    __shared__ float4 s_cm_int[kFillBlock];
    __shared__ float4 s_bb_int[kFillBlock];
    __shared__ float4 s_cm_leaf[kFillBlock];
    __shared__ uint32_t s_arrive[kFillBlock];
    __shared__ uint32_t s_done[kFillBlock];

    const uint32_t j = threadIdx.x;
    const uint32_t base = blockIdx.x * kFillBlock;
    const uint32_t nint = (uint32_t)(n - 1);  // internal nodes
    const uint32_t last = base + kFillBlock - 1 < (uint32_t)(n - 1) ? base + kFillBlock - 1
                                                                   : (uint32_t)(n - 1);
    const uint32_t t = base + j;
    s_arrive[j] = 0u;
    s_done[j] = 0u;
    if (t < nint) {
        s_child[j] = tb.child[t];
        s_range[j] = tb.range[t];
        s_par_int[j] = tb.parent[t];
    }
    float4 C, B;
    if (t < (uint64_t)n) {
        s_par_leaf[j] = n > 1 ? tb.parent[nint + t] : kNoNode;
        const uint32_t b = idx[t];
        const float2 p = pos[b];
        // lbvh.wgsl:179-187
        C = make_float4(p.x, p.y, mass[b], 0.0f);
        B = make_float4(p.x, p.y, p.x, p.y);
        s_cm_leaf[j] = C;
        tb.leaf[t] = C;
    }
    __syncthreads();

    // lbvh.wgsl:191-232, restricted to the nodes inside this block.
    if (t < (uint64_t)n && n > 1) {
        uint32_t cur = nint + t;
        uint32_t pw = s_par_leaf[j];
        for (int level = 0; level < 128; ++level) {
            const uint32_t par = pw & kParentMask;
            if (par >= nint) break;  // corrupt parent word
            const uint32_t lp = par - base;
            if (par < base || par > last || s_range[lp].x < base || s_range[lp].y > last) {
                const uint32_t sb = base / kSuper;
                const uint32_t slot = atomicAdd(&tb.counts[sb], 1u);
                tb.frontier[(size_t)sb * kSuper + slot] = cur;
                break;
            }
            const uint32_t prev = __hip_atomic_fetch_add(&s_arrive[lp], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prev == 0u) break;
            const uint2 ch = s_child[lp];
            float4 L, LB, R, RB;
            if (ch.x >= nint) {
                L = s_cm_leaf[ch.x - nint - base];
                LB = make_float4(L.x, L.y, L.x, L.y);
            } else {
                L = s_cm_int[ch.x - base];
                LB = s_bb_int[ch.x - base];
            }
            if (ch.y >= nint) {
                R = s_cm_leaf[ch.y - nint - base];
                RB = make_float4(R.x, R.y, R.x, R.y);
            } else {
                R = s_cm_int[ch.y - base];
                RB = s_bb_int[ch.y - base];
            }
            merge_values(L, LB, R, RB, C, B);
            s_cm_int[lp] = C;
            s_bb_int[lp] = B;
            s_done[lp] = 1u;
            cur = par;
            if (cur == 0u) break;
            pw = s_par_int[lp];
        }
    }
    __syncthreads();

    // Write the block's completed nodes where their readers look for them.
    if (t < (uint64_t)n) {
        float4 c = s_cm_leaf[j];
        c.w = kLeafLength;
        if (n == 1)
            publish(tb, 0u, 0u, c, make_float4(c.x, c.y, c.x, c.y));
        else
            publish(tb, nint + t, s_par_leaf[j], c, make_float4(c.x, c.y, c.x, c.y));
    }
    if (t < nint && s_done[j]) publish(tb, t, t ? s_par_int[j] : 0u, s_cm_int[j], s_bb_int[j]);
}

// Climb from `cur` (its data already published) while the parents stay in
// [lo, hi]; a node whose parent leaves that range goes to out[atomic(*out_n)].
__device__ __forceinline__ void climb_wg(const TreeBufs &tb, uint32_t nint, uint32_t cur,
                                         uint32_t lo, uint32_t hi, uint32_t *out,
                                         uint32_t *out_n) {
    for (int level = 0; level < 128; ++level) {
        const uint32_t par = tb.parent[cur] & kParentMask;
        if (par >= nint) break;
        const uint2 r = tb.range[par];
        if (r.x < lo || r.y > hi) {
            out[atomicAdd(out_n, 1u)] = cur;
            break;
        }
        const uint32_t prev = __hip_atomic_fetch_add(&tb.status[par], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prev == 0u) break;
        float4 C, B;
        merge_children(tb, par, C, B);
        cur = par;
        publish(tb, cur, cur ? tb.parent[cur] : 0u, C, B);
        if (cur == 0u) break;
    }
}

// Phase 2: workgroup sb owns super-block sb's leaves.
__global__ __launch_bounds__(kFillBlock) void fill_mid_kernel(int64_t n, TreeBufs tb,
                                                              uint32_t num_super) {
    const uint32_t sb = blockIdx.x;
    const uint32_t nint = (uint32_t)(n - 1);
    const uint32_t lo = sb * (uint32_t)kSuper;
    const uint32_t hi = lo + kSuper - 1 < nint ? lo + kSuper - 1 : nint;
    const uint32_t count = tb.counts[sb];
    for (uint32_t k = threadIdx.x; k < count; k += kFillBlock)
        climb_wg(tb, nint, tb.frontier[(size_t)sb * kSuper + k], lo, hi, tb.top,
                 &tb.counts[num_super]);
}

// Phase 3: one workgroup finishes the tree, then clears the hand-off counts.
__global__ __launch_bounds__(kFillBlock) void fill_top_kernel(int64_t n, TreeBufs tb,
                                                              uint32_t num_super) {
    const uint32_t nint = (uint32_t)(n - 1);
    const uint32_t count = tb.counts[num_super];
    for (uint32_t k = threadIdx.x; k < count; k += kFillBlock)
        climb_wg(tb, nint, tb.top[k], 0u, nint, tb.top, &tb.counts[num_super]);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k <= num_super; k += kFillBlock) tb.counts[k] = 0u;
}

}  // namespace

void launch_build_u32(const uint32_t *keys, int64_t n, const TreeBufs &t, hipStream_t s) {
    if (n < 2) return;
    build_u32_kernel<<<(unsigned)((n - 1 + 255) / 256), 256, 0, s>>>(keys, n, t);
}

void launch_build_u64(const uint64_t *keys, int64_t n, const int64_t *sim_off, const TreeBufs &t,
                      uint32_t *sim_root, hipStream_t s) {
    if (n < 2) return;
    build_u64_kernel<<<(unsigned)((n - 1 + 255) / 256), 256, 0, s>>>(keys, n, sim_off, t,
                                                                     sim_root);
}

void launch_fill(const uint32_t *idx, const float *mass, const float2 *pos, int64_t n,
                 const TreeBufs &t, hipStream_t s) {
    if (n < 1) return;
    const unsigned blocks = (unsigned)((n + kFillBlock - 1) / kFillBlock);
    fill_local_kernel<<<blocks, kFillBlock, 0, s>>>(idx, mass, pos, n, t);
    if (n > kFillBlock) {
        const uint32_t num_super = (uint32_t)((n + kSuper - 1) / kSuper);
        fill_mid_kernel<<<num_super, kFillBlock, 0, s>>>(n, t, num_super);
        fill_top_kernel<<<1, kFillBlock, 0, s>>>(n, t, num_super);
    }
}

int64_t fill_frontier_words(int64_t cap) { return ((cap + kSuper - 1) / kSuper) * kSuper; }
int64_t fill_count_words(int64_t cap) { return (cap + kSuper - 1) / kSuper + 1; }

}  // namespace nbs
