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
//   every thread at every level it costs milliseconds.  The climb is split
//   so that almost all of it is ordered at workgroup scope instead:
//     phase 1 (fill_local_kernel): workgroup b owns sorted leaves
//       [b*1024, b*1024+1023] and completes every internal node whose leaf
//       range lies inside that block, with the arrival counters in LDS.  A
//       thread that completes a node whose parent spans blocks appends the
//       node to a frontier list.
//     phase 2 (fill_top_kernel): a few workgroups climb from the frontier
//       nodes through the block-spanning nodes (a few per block boundary)
//       with agent-scope counters; the kernel boundary makes phase 1's nodes
//       visible to them.
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

// lbvh.wgsl:205-228: node p from its two complete children, read from
// p's child-pair slots.
__device__ __forceinline__ void merge_children(const TreeBufs &tb, uint32_t p, float4 &C,
                                               float4 &B) {
    const float4 L = tb.pcm[2 * p], R = tb.pcm[2 * p + 1];
    const float4 LB = tb.paabb[2 * p], RB = tb.paabb[2 * p + 1];
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

// Store a completed node's data where its readers look for it: the child
// slot of its parent (parent word `pw`), or the root record.
__device__ __forceinline__ void publish(const TreeBufs &tb, uint32_t node, uint32_t pw,
                                        const float4 &C, const float4 &B) {
    if (node == 0u) {
        tb.root[0] = C;
        tb.root[1] = B;
        return;
    }
    const uint32_t slot = 2u * (pw & kParentMask) + (pw >> 31);
    tb.pcm[slot] = C;
    tb.paabb[slot] = B;
}

__global__ __launch_bounds__(kFillBlock) void fill_local_kernel(const uint32_t *__restrict__ idx,
                                                                const float *__restrict__ mass,
                                                                const float2 *__restrict__ pos,
                                                                int64_t n, TreeBufs tb) {
    __shared__ uint32_t s_arrive[kFillBlock];
    s_arrive[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kFillBlock;
    const uint32_t last = (uint32_t)((int64_t)base + kFillBlock - 1 < n - 1 ? base + kFillBlock - 1
                                                                             : n - 1);
    const uint32_t t = base + threadIdx.x;
    if (t >= (uint64_t)n) return;
    const uint32_t nint = (uint32_t)(n - 1);  // internal nodes
    uint32_t cur = nint + t;
    const uint32_t b = idx[t];
    const float2 p = pos[b];
    // lbvh.wgsl:179-187
    float4 C = make_float4(p.x, p.y, mass[b], 0.0f);
    float4 B = make_float4(p.x, p.y, p.x, p.y);
    tb.leaf[t] = C;
    if (n == 1) {
        publish(tb, 0u, 0u, C, B);
        return;
    }
    // lbvh.wgsl:191-232, restricted to nodes inside this block's leaf range.
    // A valid tree reaches the root in <= 96 levels; the level bound and the
    // index check only guard against a corrupt parent array.
    for (int level = 0; level < 128; ++level) {
        const uint32_t pw = tb.parent[cur];
        const uint32_t par = pw & kParentMask;
        if (par >= nint) break;
        publish(tb, cur, pw, C, B);
        const uint2 r = tb.range[par];
        if (r.x < base || r.y > last) {
            const uint32_t slot = atomicAdd(tb.frontier_count, 1u);
            tb.frontier[slot] = cur;
            break;
        }
        const uint32_t prev = __hip_atomic_fetch_add(&s_arrive[par - base], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prev == 0u) break;
        merge_children(tb, par, C, B);
        cur = par;
        if (cur == 0u) {
            publish(tb, 0u, 0u, C, B);
            break;
        }
    }
}

// Phase 2: climb from the frontier through the block-spanning nodes (a few
// per block boundary).  Threads of several workgroups (on any XCD) meet at
// these nodes, so their counters are agent-scope acquire-release; with only
// the frontier's few thousand threads doing so, the L2 write-back and
// invalidate they imply stay cheap.  Each frontier node's data is already
// in its parent's child slot.
constexpr int kTopBlocks = 32;
constexpr int kTopThreads = 256;

__global__ __launch_bounds__(kTopThreads) void fill_top_kernel(int64_t n, TreeBufs tb) {
    const uint32_t count = *tb.frontier_count;
    const uint32_t nint = (uint32_t)(n - 1);
    for (uint32_t k = blockIdx.x * kTopThreads + threadIdx.x; k < count;
         k += kTopBlocks * kTopThreads) {
        uint32_t cur = tb.frontier[k];
        for (int level = 0; level < 128; ++level) {
            const uint32_t par = tb.parent[cur] & kParentMask;
            if (par >= nint) break;
            const uint32_t prev = __hip_atomic_fetch_add(&tb.status[par], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            if (prev == 0u) break;
            float4 C, B;
            merge_children(tb, par, C, B);
            cur = par;
            publish(tb, cur, cur ? tb.parent[cur] : 0u, C, B);
            if (cur == 0u) break;
        }
    }
}

// Reset the frontier counter for the next step (after phase 2 read it).
__global__ void fill_reset_kernel(uint32_t *count) { *count = 0u; }

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
    fill_local_kernel<<<(unsigned)((n + kFillBlock - 1) / kFillBlock), kFillBlock, 0, s>>>(
        idx, mass, pos, n, t);
    if (n > kFillBlock) {
        fill_top_kernel<<<kTopBlocks, kTopThreads, 0, s>>>(n, t);
        fill_reset_kernel<<<1, 1, 0, s>>>(t.frontier_count);
    }
}

}  // namespace nbs
