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
//   every thread at every level it costs milliseconds.  The fill is split in
//   phases whose ordering is at workgroup scope:
//     phase 1 (fill_local_kernel): workgroup b owns sorted leaves
//       [b*1024, b*1024+1023].  It loads its slice of the child and parent
//       arrays into LDS and completes every node whose leaf range lies in
//       the block (the build flags their children's parent words with
//       kInBlock) with LDS-resident node data and LDS arrival counters, then
//       writes all completed nodes out cooperatively.
//     phase 2 (fill_mid_kernel): one workgroup per super-block of 32 blocks
//       finds, among its node indices, the nodes whose range lies in the
//       super-block across a block boundary, and completes them by LDS
//       climbs; it also lists the top nodes (range across super-blocks).
//     phase 3 (the last fill_mid workgroup to finish): completes the top.
//   A kernel boundary orders phases 1 and 2; one device-scope fence per
//   phase-2 workgroup orders phases 2 and 3.
//   The merge formula does not depend on which child arrives last, so the
//   result is deterministic and equals a sequential post-order fill.  The
//   reference's random leaf rotation (`offset`, lbvh.wgsl:171) is therefore
//   unnecessary and not applied.
#include "nbody_internal.h"

namespace nbs {

namespace {

// Sorted keys as the build reads them: the workgroup's own keys and a halo
// of kBuildHalo keys on either side from LDS, farther keys from memory.  The
// searches of a node probe its neighbourhood first, so almost every probe
// is an LDS read: only nodes with wide leaf ranges go to memory, for their
// first (farthest) probes.
constexpr int kBuildThreads = 256;
constexpr int kBuildPerThread = 1;  // nodes per thread (2: all wavefronts resident at once, measured slower)
constexpr int kBuildNodes = kBuildThreads * kBuildPerThread;  // nodes per workgroup
constexpr int kBuildHalo = 1024;
constexpr int kBuildWindow = kBuildNodes + 2 * kBuildHalo;
// A global-memory read the compiler cannot merge with another load; it
// waits for its data in the same asm block.
__device__ __forceinline__ uint32_t global_read(const uint32_t *p) {
    uint32_t v;
    asm volatile("global_load_dword %0, %1, off\n\ts_waitcnt vmcnt(0)" : "=v"(v) : "v"(p) : "memory");
    return v;
}
__device__ __forceinline__ uint64_t global_read(const uint64_t *p) {
    uint64_t v;
    asm volatile("global_load_dwordx2 %0, %1, off\n\ts_waitcnt vmcnt(0)" : "=v"(v) : "v"(p) : "memory");
    return v;
}

// Indices are 32-bit: positions stay below 2^30 (the parent words' limit).
template <typename K>
struct KeyWindow {
    const K *__restrict__ g;  // all keys
    const K *s;               // LDS copy of keys [lo, lo + len)
    int lo, len;
    // an LDS read for every lane (position 0 outside the window), then a
    // memory read for the lanes outside it, written so that the compiler
    // cannot merge the two into one load from a selected address (every
    // probe would become a flat load)
    __device__ __forceinline__ K operator[](int j) const {
        const bool in = (uint32_t)(j - lo) < (uint32_t)len;
        K v = s[in ? j - lo : 0];
        if (!in) v = global_read(g + j);
        return v;
    }
};
// Load the window of the workgroup's nodes [b0, b0 + kBuildNodes) into s_keys.
template <typename K>
__device__ __forceinline__ KeyWindow<K> load_window(const K *__restrict__ keys, int n,
                                                    K *s_keys) {
    const int b0 = (int)blockIdx.x * kBuildNodes;
    const int lo = b0 > kBuildHalo ? b0 - kBuildHalo : 0;
    const int hi = b0 + kBuildNodes + kBuildHalo < n ? b0 + kBuildNodes + kBuildHalo : n;
    for (int j = lo + (int)threadIdx.x; j < hi; j += kBuildThreads) s_keys[j - lo] = keys[j];
    __syncthreads();
    return KeyWindow<K>{keys, s_keys, lo, hi - lo};
}

// lbvh.wgsl:72-83: the common prefix length of sorted keys i and j, -1 for
// j outside [0, n); equal keys compare their positions instead.  A node's
// searches compare every probe with its own key, read once (at(i)).
struct DeltaU32 {
    KeyWindow<uint32_t> k;
    int n;
    struct Node {
        const DeltaU32 &d;
        uint32_t ki;
        int i;
        __device__ __forceinline__ int operator()(int j) const {
            if ((uint32_t)j >= (uint32_t)d.n) return -1;
            const uint32_t x = ki ^ d.k[j];
            const int by_key = __clz((int)x), by_pos = 32 + __clz(i ^ j);
            return x ? by_key : by_pos;
        }
    };
    __device__ __forceinline__ Node at(int i) const { return Node{*this, k[i], i}; }
};

// Ensemble keys carry the sim id above bit 30.  Equal keys share a sim and
// are tie-broken on sim-local sorted positions, so every sim's sub-tree has
// the shape of the tree a plain run of that sim builds.
struct DeltaU64 {
    KeyWindow<uint64_t> k;
    const int64_t *__restrict__ sim_off;
    int n;
    struct Node {
        const DeltaU64 &d;
        uint64_t ki;
        int i;
        __device__ __forceinline__ int operator()(int j) const {
            if ((uint32_t)j >= (uint32_t)d.n) return -1;
            const uint64_t b = d.k[j];
            if (ki == b) {
                const int off = (int)d.sim_off[ki >> kSimShift];
                return 64 + __clz((i - off) ^ (j - off));
            }
            return __clzll((long long)(ki ^ b));
        }
    };
    __device__ __forceinline__ Node at(int i) const { return Node{*this, k[i], i}; }
};

struct Split {
    int lo, hi, gamma;
    uint32_t lc, rc;
};

// Fill hierarchy: phase-1 blocks of 1024 sorted leaves, 32 blocks to a
// super-block.  A node belongs to the phase whose unit holds its whole leaf
// range: a block (phase 1), a super-block (phase 2), else the top (phase 3).
constexpr int kBlockShift = 10;
constexpr int kSuperShift = 15;
constexpr int kFillBlock = 1 << kBlockShift;
constexpr int kSuper = 1 << kSuperShift;                         // leaves per super-block
// The fill's upper levels: the phase-2 workgroup of super-block s finds
// its phase-2 nodes (and the top nodes) by their leaf ranges among node
// indices [s * kSuper, (s+1) * kSuper): a node's index is an end of its
// range, so a node whose range lies in super-block s has its index there.
constexpr uint32_t kLdsList = 1024;  // longest list the fill resolves in LDS
constexpr uint32_t kScan = kSuper / 1024;  // node indices per thread of the scan
__host__ __device__ inline uint32_t num_supers(int64_t n) {
    return (uint32_t)((n + kSuper - 1) / kSuper);
}
// The top nodes (ranges across super-blocks) whose index lies in
// super-block s are the nodes whose range holds its left or right boundary:
// two root-ward chains, at most 2 x the tree depth (< 2 x 97 for 38-bit
// keys tie-broken on positions below 2^30).  Each super-block lists them in
// its own region of kTopPer words of tb.lists, so no workgroup waits on
// another to reserve space; tb.slot holds each top node's list position.
constexpr uint32_t kTopPer = 256;
// Counter words (tb.counts): [0] phase-2 done count, [2 + s] number of top
// nodes of super-block s.

__device__ __forceinline__ bool spans(uint2 r, int shift) { return (r.x >> shift) != (r.y >> shift); }

// kInBlock when the node's leaf range lies in one phase-1 block.
__device__ __forceinline__ uint32_t in_block(const Split &sp) {
    return ((uint32_t)sp.lo >> kBlockShift) == ((uint32_t)sp.hi >> kBlockShift) ? kInBlock : 0u;
}

template <typename Delta>
__device__ __forceinline__ Split karras_node(const Delta &delta, int i, int n) {
    // lbvh.wgsl:101-152; probe i + d * x is written as a select on d
    const auto dl = delta.at(i);
    const int d = dl(i + 1) > dl(i - 1) ? 1 : -1;
    const int delta_min = dl(i - d);
    int l_max = 2;
    while (dl(d > 0 ? i + l_max : i - l_max) > delta_min) l_max *= 2;
    int l = 0;
    for (int t = l_max >> 1; t >= 1; t >>= 1)
        if (dl(d > 0 ? i + l + t : i - l - t) > delta_min) l += t;
    const int j = d > 0 ? i + l : i - l;
    const int delta_node = dl(j);
    int s = 0;
    int t = (l + 1) >> 1;
    for (;;) {
        if (dl(d > 0 ? i + s + t : i - s - t) > delta_node) s += t;
        if (t == 1) break;
        t = (t + 1) >> 1;
    }
    Split r;
    r.gamma = d > 0 ? i + s : i - s - 1;
    r.lo = i < j ? i : j;
    r.hi = i < j ? j : i;
    r.lc = (r.lo == r.gamma) ? (uint32_t)(n - 1 + r.gamma) : (uint32_t)r.gamma;
    r.rc = (r.hi == r.gamma + 1) ? (uint32_t)(n - 1 + r.gamma + 1) : (uint32_t)(r.gamma + 1);
    return r;
}

// Zero the sort's scratch for the next step (every sort is followed by a
// build, which runs after the sort has finished with it), and the walk's
// bounding-box accumulators (the encode has read them; the walk that
// follows refills them).
__device__ __forceinline__ void clear_words(uint32_t *__restrict__ p, int64_t words,
                                            uint32_t *__restrict__ slots) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
    if (slots && blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < (uint32_t)kBBoxSlots; k += blockDim.x)
            reinterpret_cast<uint4 *>(slots)[k] = make_uint4(0u, 0u, 0u, 0u);
}

// Leaf record {x, y, mass, 0} of sorted position i (lbvh.wgsl:179-187).
// The sort's last pass writes them (its gathers overlap its ranking); a
// single body, which has no sort, is gathered by the fill.
__device__ __forceinline__ float4 gather_leaf(const TreeBufs &tb, const uint32_t *__restrict__ idx,
                                              const float *__restrict__ mass,
                                              const float2 *__restrict__ pos, int i) {
    const uint32_t b = idx[i];
    const float2 p = pos[b];
    const float4 c = make_float4(p.x, p.y, mass[b], 0.0f);
    tb.leaf[i] = c;
    return c;
}

__global__ __launch_bounds__(kBuildThreads) void build_u32_kernel(
    const uint32_t *__restrict__ keys, int64_t n, TreeBufs tb, uint32_t *__restrict__ clear,
    int64_t clear_n, uint32_t *__restrict__ slots) {
    __shared__ uint32_t s_keys[kBuildWindow];
    clear_words(clear, clear_n, slots);
    const int n32 = (int)n;
    const DeltaU32 delta{load_window(keys, n32, s_keys), n32};
    for (int r = 0; r < kBuildPerThread; ++r) {
        const int i = (int)blockIdx.x * kBuildNodes + r * kBuildThreads + (int)threadIdx.x;
        if (i >= n32 - 1) return;
        const Split sp = karras_node(delta, i, n32);
        tb.child[i] = make_uint2(sp.lc, sp.rc);
        // the walk reads the indices from the pair record
        *reinterpret_cast<uint2 *>(tb.pcm + kPairStride * (size_t)i + 2) = make_uint2(sp.lc, sp.rc);
        tb.range[i] = make_uint2((uint32_t)sp.lo, (uint32_t)sp.hi);
        // lbvh.wgsl:155-161, with the fill's in-block flag
        const uint32_t pw = (uint32_t)i | in_block(sp);
        tb.parent[sp.lc] = pw;
        tb.parent[sp.rc] = pw | kRightChild;
        if (i == 0) tb.parent[0] = kNoNode;
        tb.status[i] = 0u;  // lbvh.wgsl:164
    }
}

__global__ __launch_bounds__(kBuildThreads) void build_u64_kernel(const uint64_t *__restrict__ keys,
                                                        int64_t n,
                                                        const int64_t *__restrict__ sim_off,
                                                        TreeBufs tb,
                                                        uint32_t *__restrict__ sim_root,
                                                        uint32_t *__restrict__ clear,
                                                        int64_t clear_n,
                                                        uint32_t *__restrict__ slots) {
    __shared__ uint64_t s_keys[kBuildWindow];
    clear_words(clear, clear_n, slots);
    const int n32 = (int)n;
    const DeltaU64 delta{load_window(keys, n32, s_keys), sim_off, n32};
    for (int r = 0; r < kBuildPerThread; ++r) {
        const int i = (int)blockIdx.x * kBuildNodes + r * kBuildThreads + (int)threadIdx.x;
        if (i >= n32 - 1) return;
        const Split sp = karras_node(delta, i, n32);
        tb.child[i] = make_uint2(sp.lc, sp.rc);
        // the walk reads the indices from the pair record
        *reinterpret_cast<uint2 *>(tb.pcm + kPairStride * (size_t)i + 2) = make_uint2(sp.lc, sp.rc);
        tb.range[i] = make_uint2((uint32_t)sp.lo, (uint32_t)sp.hi);
        const uint32_t pw = (uint32_t)i | in_block(sp);
        tb.parent[sp.lc] = pw;
        tb.parent[sp.rc] = pw | kRightChild;
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
}


// lbvh.wgsl:205-228: a node from its two complete children.
__device__ __forceinline__ void merge_values(const float4 &L, const float4 &LB, const float4 &R,
                                             const float4 &RB, float4 &C, float4 &B) {
    B.x = fminf(LB.x, RB.x);
    B.y = fminf(LB.y, RB.y);
    B.z = fmaxf(LB.z, RB.z);
    B.w = fmaxf(LB.w, RB.w);
    C.z = L.z + R.z;
    // both quotients, then a select: a branch here would hold back the
    // children's loads until the mass test
    const float qx = (L.x * L.z + R.x * R.z) / C.z;
    const float qy = (L.y * L.z + R.y * R.z) / C.z;
    C.x = C.z > 0.0f ? qx : 0.0f;
    C.y = C.z > 0.0f ? qy : 0.0f;
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


__global__ __launch_bounds__(kFillBlock) void fill_local_kernel(const uint32_t *__restrict__ idx,
                                                                const float *__restrict__ mass,
                                                                const float2 *__restrict__ pos,
                                                                int64_t n, TreeBufs tb) {
    __shared__ uint2 s_child[kFillBlock];  // internal node base+j
    __shared__ uint32_t s_par_int[kFillBlock];
    // {cm, mass, length}: internal node base+j at j, leaf base+j at
    // kFillBlock+j; AABBs of the internal nodes (a leaf's is its position)
    __shared__ float4 s_cm[2 * kFillBlock];
    __shared__ float4 s_bb[kFillBlock];
    __shared__ uint32_t s_arrive[kFillBlock];  // 2: the node is complete

    const uint32_t j = threadIdx.x;
    const uint32_t base = blockIdx.x * kFillBlock;
    const uint32_t nint = (uint32_t)(n - 1);  // internal nodes
    const uint32_t t = base + j;
    s_arrive[j] = 0u;
    if (t < nint) {
        s_child[j] = tb.child[t];
        s_par_int[j] = tb.parent[t];
    }
    uint32_t pw_leaf = kNoNode;
    float4 C = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (t < (uint64_t)n) {
        pw_leaf = n > 1 ? tb.parent[nint + t] : kNoNode;
        // the build gathered the leaf records; a single body has no build
        C = n == 1 ? gather_leaf(tb, idx, mass, pos, 0) : tb.leaf[t];
        s_cm[kFillBlock + j] = C;
    }
    __syncthreads();

    // lbvh.wgsl:191-232, restricted to the nodes inside this block: a node
    // whose parent word lacks kInBlock is a top of the block's forest.  The
    // climbing thread keeps its node in registers; the child that arrives
    // first left its record in LDS, so a level is the arrival count (its
    // wait also covers the parent's child pair and parent word, read
    // beforehand), one read of the sibling and the merge.
    if (t < (uint64_t)n && n > 1) {
        float4 B = make_float4(C.x, C.y, C.x, C.y);
        uint32_t pw = pw_leaf;
        uint32_t lp = (pw & kParentMask) - base;
        bool up = (pw & kInBlock) && lp < (uint32_t)kFillBlock;
        uint2 ch = make_uint2(0u, 0u);
        uint32_t ppw = 0u;
        if (up) {
            ch = s_child[lp];
            ppw = s_par_int[lp];
        }
        for (int level = 0; up && level < 128; ++level) {
            const uint32_t prev = __hip_atomic_fetch_add(&s_arrive[lp], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prev == 0u) break;
            const bool right = (pw >> 31) != 0u;
            const uint32_t sib = right ? ch.x : ch.y;
            const bool sleaf = sib >= nint;
            const uint32_t si = sleaf ? sib - nint - base + (uint32_t)kFillBlock : sib - base;
            // one read from a selected LDS address (a leaf's record holds
            // its AABB corner), so both reads issue together
            const float4 S = s_cm[si];
            float4 SR = *(sleaf ? &s_cm[si] : &s_bb[si]);
            // keep the read whole (else .zw become a second, conditional read)
            asm volatile("" : "+v"(SR.z), "+v"(SR.w));
            const float4 SB = sleaf ? make_float4(SR.x, SR.y, SR.x, SR.y) : SR;
            float4 NC, NB;
            merge_values(right ? S : C, right ? SB : B, right ? C : S, right ? B : SB, NC, NB);
            C = NC;
            B = NB;
            s_cm[lp] = C;
            s_bb[lp] = B;
            const uint32_t par = lp + base;
            pw = ppw;
            lp = (pw & kParentMask) - base;
            up = par != 0u && (pw & kInBlock) && lp < (uint32_t)kFillBlock;
            if (up) {
                ch = s_child[lp];
                ppw = s_par_int[lp];
            }
        }
    }
    __syncthreads();

    // Write the block's completed nodes where their readers look for them.
    if (t < (uint64_t)n) {
        float4 c = s_cm[kFillBlock + j];
        c.w = kLeafLength;
        if (n == 1)
            publish(tb, 0u, 0u, c, make_float4(c.x, c.y, c.x, c.y));
        else
            publish(tb, nint + t, pw_leaf, c, make_float4(c.x, c.y, c.x, c.y));
    }
    if (t < nint && s_arrive[j] == 2u)
        publish(tb, t, t ? s_par_int[j] : 0u, s_cm[j], s_bb[j]);
}

// Climb from `cur` (its data already published) while the parents' leaf
// ranges stay in [lo, hi] (global arrival counters, workgroup scope).  A
// parent's range and parent word come from the build (an earlier kernel),
// so both are read together, before the arrival counter: a level costs three
// dependent trips to memory (those reads, the counter, the children's data).
__device__ __forceinline__ void climb_wg(const TreeBufs &tb, uint32_t nint, uint32_t cur,
                                         uint32_t lo, uint32_t hi) {
    uint32_t pw = tb.parent[cur];
    for (int level = 0; level < 128; ++level) {
        const uint32_t par = pw & kParentMask;
        if (par >= nint) break;
        const uint2 r = tb.range[par];
        const uint32_t ppw = par ? tb.parent[par] : 0u;
        if (r.x < lo || r.y > hi) break;
        const uint32_t prev = __hip_atomic_fetch_add(&tb.status[par], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prev == 0u) break;
        float4 C, B;
        merge_children(tb, par, C, B);
        publish(tb, par, ppw, C, B);
        cur = par;
        pw = ppw;
        if (cur == 0u) break;
    }
}

// Exclusive prefix sum over the kFillBlock threads of a workgroup; *total
// receives the sum.  sh: kFillBlock / 64 + 1 words of LDS.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kFillBlock / 64; ++w) {
        const uint32_t t = sh[w];
        before += w < wid ? t : 0u;
        all += t;
    }
    *total = all;
    return before + x - v;
}

// One upper-level node list in LDS: entry e is node node[e] with parent
// word pw[e], its parent's entry par[e] (kNone: the parent is not in the
// list), the number of its children still pending, and its children's
// {cm, aabb} records (left, right).
constexpr uint32_t kNone = 0xFFFFFFFFu;
struct ListLds {
    uint32_t node[kLdsList], pw[kLdsList], par[kLdsList], pend[kLdsList];
    float4 in[kLdsList][4];
    float4 out[kLdsList][2];  // the entry's own {cm, aabb} once complete
};

// Index of `x` in the sorted node[0, m), or kNone.
__device__ __forceinline__ uint32_t find_node(const ListLds &ls, uint32_t m, uint32_t x) {
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ls.node[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < m && ls.node[lo] == x ? lo : kNone;
}

// Resolve a list set up in ls (m <= kLdsList entries; called by the whole
// workgroup).  The entries with no pending child start a climb in LDS: merge
// the children (lbvh.wgsl:205-228), keep the node in its entry, hand it to
// the parent entry's input and count the parent's arrival (workgroup-scope
// acquire/release); the child that arrives last climbs on, the other stops.
// The climbing thread keeps its node in registers: a level is the merge,
// the LDS writes with the parent's fixed words read alongside, the arrival
// count and one read of the sibling's record.  The chains of the upper
// levels are long (tens of levels), so the nodes are published to memory
// only afterwards, by all threads together.
__device__ __forceinline__ void resolve_list(const TreeBufs &tb, ListLds &ls, uint32_t m) {
    const uint32_t e = threadIdx.x;
    if (e < m && ls.pend[e] == 0u) {
        uint32_t cur = e;
        float4 L = ls.in[cur][0], LB = ls.in[cur][1], R = ls.in[cur][2], RB = ls.in[cur][3];
        uint32_t pw = ls.pw[cur], q = ls.par[cur];
        for (int level = 0; level < 128; ++level) {
            float4 C, B;
            merge_values(L, LB, R, RB, C, B);
            ls.out[cur][0] = C;
            ls.out[cur][1] = B;
            if (q == kNone) break;
            const uint32_t side = pw >> 31;
            ls.in[q][2 * side] = C;
            ls.in[q][2 * side + 1] = B;
            // the parent's own words are fixed: read before the arrival
            const uint32_t qpw = ls.pw[q], qq = ls.par[q];
            if (__hip_atomic_fetch_sub(&ls.pend[q], 1u, __ATOMIC_ACQ_REL,
                                       __HIP_MEMORY_SCOPE_WORKGROUP) != 1u)
                break;  // the sibling is still pending: it climbs on
            // the sibling's record, there before its arrival (or preloaded)
            const float4 S = ls.in[q][2 * (side ^ 1u)], SB = ls.in[q][2 * (side ^ 1u) + 1];
            L = side ? S : C;
            LB = side ? SB : B;
            R = side ? C : S;
            RB = side ? B : SB;
            cur = q;
            pw = qpw;
            q = qq;
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < m; k += kFillBlock)
        publish(tb, ls.node[k], ls.pw[k], ls.out[k][0], ls.out[k][1]);
}

// Phases 2 and 3.  Workgroup sb scans the leaf ranges of node indices
// [sb * kSuper, (sb+1) * kSuper): its phase-2 nodes (range inside the
// super-block, across a block boundary) become an LDS list in index order,
// its top nodes (range across super-blocks) go to the global top list.  The
// phase-2 list is resolved in LDS (or, longer than lds_limit, by climbing
// from the phase-2 nodes' other children).  The workgroup that finishes last (a
// device-scope fence orders each workgroup's writes before its count)
// resolves the top list the same way (or climbs from the top nodes'
// other children) and clears the counters.
__global__ __launch_bounds__(kFillBlock) void fill_mid_kernel(int64_t n, TreeBufs tb,
                                                              uint32_t num_super,
                                                              uint32_t lds_limit) {
    __shared__ ListLds ls;
    __shared__ uint32_t sh[kFillBlock / 64 + 1];
    __shared__ uint32_t s_cnt[kScan * 16];
    __shared__ uint32_t s_pre[kFillBlock];  // top: list offsets of the super-blocks
    __shared__ uint32_t s_last;
    const uint32_t t = threadIdx.x;
    const uint32_t sb = blockIdx.x;
    const uint32_t nint = (uint32_t)(n - 1);
    uint32_t *top_list = tb.lists;

    // scan: node index base_i + k * kFillBlock + t (coalesced), every load
    // issued before any is used (indices clamped into the array)
    const uint32_t lane = t & 63, w = t >> 6;
    const uint32_t base_i = sb * (uint32_t)kSuper;
    uint2 r[kScan];
#pragma unroll
    for (uint32_t k = 0; k < kScan; ++k) r[k] = tb.range[min(base_i + k * kFillBlock + t, nint - 1u)];
    uint32_t bits2 = 0, bits3 = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScan; ++k) {
        if (base_i + k * kFillBlock + t < nint && spans(r[k], kBlockShift)) {
            if (spans(r[k], kSuperShift))
                bits3 |= 1u << k;
            else
                bits2 |= 1u << k;
        }
    }
    // list positions in index order: counts per (k, wave), phase-2 in the
    // low and top in the high half-word, one scan
#pragma unroll
    for (uint32_t k = 0; k < kScan; ++k) {
        const uint64_t b2 = __ballot((bits2 >> k) & 1u), b3 = __ballot((bits3 >> k) & 1u);
        if (lane == 0) s_cnt[k * 16 + w] = __popcll(b2) | (uint32_t)__popcll(b3) << 16;
    }
    __syncthreads();
    const uint32_t v = t < kScan * 16 ? s_cnt[t] : 0u;
    uint32_t total;
    const uint32_t excl = block_scan(v, sh, &total);
    if (t < kScan * 16) s_cnt[t] = excl;
    const uint32_t m2 = total & 0xFFFFu, n3 = total >> 16;
    if (t == 0) tb.counts[2 + sb] = n3;
    const bool lds2 = m2 <= lds_limit;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kScan; ++k) {
        const uint64_t b2 = __ballot((bits2 >> k) & 1u), b3 = __ballot((bits3 >> k) & 1u);
        const uint32_t o = s_cnt[k * 16 + w];
        const uint32_t i = base_i + k * kFillBlock + t;
        const uint32_t below = __builtin_amdgcn_mbcnt_lo((uint32_t)b2, 0u);
        if ((bits2 >> k) & 1u) {
            if (lds2)
                ls.node[(o & 0xFFFFu) +
                        __builtin_amdgcn_mbcnt_hi((uint32_t)(b2 >> 32), below)] = i;
        }
        if ((bits3 >> k) & 1u) {
            const uint32_t local =
                (o >> 16) + __builtin_amdgcn_mbcnt_hi((uint32_t)(b3 >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)b3, 0u));
            const uint32_t pos = sb * kTopPer + local;  // local < kTopPer (two chains)
            top_list[pos] = i;
            tb.slot[i] = pos;
        }
    }
    __syncthreads();

    if (lds2) {
        if (t < m2) {
            const uint32_t p = ls.node[t];
            const uint32_t pw = tb.parent[p];
            const uint2 ch = tb.child[p];
            const uint32_t par = pw & kParentMask;
            // a parent spanning super-blocks is not in the list (not found)
            ls.par[t] = par < nint ? find_node(ls, m2, par) : kNone;
            uint32_t pend = 0;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t c = side ? ch.y : ch.x;
                if (c < nint && find_node(ls, m2, c) != kNone) {
                    ++pend;
                } else {
                    ls.in[t][2 * side] = tb.pcm[kPairStride * (size_t)p + side];
                    ls.in[t][2 * side + 1] = tb.paabb[2 * (size_t)p + side];
                }
            }
            ls.pw[t] = pw;
            ls.pend[t] = pend;
        }
        __syncthreads();
        resolve_list(tb, ls, m2);
    } else {
        // climb from every child of a phase-2 node that is not one itself
        const uint32_t lo = sb * (uint32_t)kSuper;
        const uint32_t hi = lo + kSuper - 1 < nint ? lo + kSuper - 1 : nint;
        for (uint32_t b = bits2; b; b &= b - 1) {
            const uint2 ch = tb.child[base_i + __builtin_ctz(b) * kFillBlock + t];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t c = side ? ch.y : ch.x;
                if (!(c < nint && spans(tb.range[c], kBlockShift))) climb_wg(tb, nint, c, lo, hi);
            }
        }
    }

    __syncthreads();
    if (t == 0) {
        __threadfence();
        s_last = atomicAdd(&tb.counts[0], 1u) == num_super - 1u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // The top list: the super-blocks' regions, compacted through their
    // counts' prefix sums (entry e lies in the last region starting at or
    // before e).
    const uint32_t S = num_super;
    uint32_t m3 = 0xFFFFFFFFu;  // more super-blocks than threads: climb instead
    if (S <= (uint32_t)kFillBlock) {
        const uint32_t c = t < S ? tb.counts[2 + t] : 0u;
        const uint32_t pre = block_scan(c, sh, &m3);
        if (t < S) s_pre[t] = pre;
        __syncthreads();
    }
    if (m3 <= lds_limit) {
        if (t < m3) {
            uint32_t lo = 0, hi = S;  // last region with s_pre <= t
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pre[mid] <= t)
                    lo = mid;
                else
                    hi = mid;
            }
            const uint32_t p = top_list[lo * kTopPer + (t - s_pre[lo])];
            const uint32_t pw = tb.parent[p];
            const uint2 ch = tb.child[p];
            const uint32_t par = pw & kParentMask;
            // every ancestor of a top node is a top node
            uint32_t q = kNone;
            if (par < nint) {
                const uint32_t ps = tb.slot[par];
                q = s_pre[ps / kTopPer] + ps % kTopPer;
            }
            ls.par[t] = q;
            ls.node[t] = p;
            uint32_t pend = 0;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t c = side ? ch.y : ch.x;
                if (c < nint && spans(tb.range[c], kSuperShift)) {
                    ++pend;
                } else {
                    ls.in[t][2 * side] = tb.pcm[kPairStride * (size_t)p + side];
                    ls.in[t][2 * side + 1] = tb.paabb[2 * (size_t)p + side];
                }
            }
            ls.pw[t] = pw;
            ls.pend[t] = pend;
        }
        __syncthreads();
        resolve_list(tb, ls, m3);
    } else {
        // climb from every child of a top node that is not a top node
        for (uint32_t k = t; k < S * kTopPer; k += kFillBlock) {
            if (k % kTopPer >= tb.counts[2 + k / kTopPer]) continue;
            const uint2 ch = tb.child[top_list[k]];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t c = side ? ch.y : ch.x;
                if (!(c < nint && spans(tb.range[c], kSuperShift))) climb_wg(tb, nint, c, 0u, nint);
            }
        }
    }
    __syncthreads();
    if (t == 0) tb.counts[0] = 0u;
}

}  // namespace

void launch_build_u32(const uint32_t *keys, int64_t n, const TreeBufs &t, uint32_t *clear,
                      int64_t clear_n, uint32_t *slots, hipStream_t s) {
    if (n < 2) return;
    build_u32_kernel<<<(unsigned)((n - 1 + kBuildNodes - 1) / kBuildNodes), kBuildThreads, 0,
                       s>>>(keys, n, t, clear, clear_n, slots);
}

void launch_build_u64(const uint64_t *keys, int64_t n, const int64_t *sim_off, const TreeBufs &t,
                      uint32_t *sim_root, uint32_t *clear, int64_t clear_n, uint32_t *slots,
                      hipStream_t s) {
    if (n < 2) return;
    build_u64_kernel<<<(unsigned)((n - 1 + kBuildNodes - 1) / kBuildNodes), kBuildThreads, 0,
                       s>>>(keys, n, sim_off, t, sim_root, clear, clear_n, slots);
}

void launch_fill(const uint32_t *idx, const float *mass, const float2 *pos, int64_t n,
                 const TreeBufs &t, bool climb, hipStream_t s) {
    if (n < 1) return;
    const unsigned blocks = (unsigned)((n + kFillBlock - 1) / kFillBlock);
    fill_local_kernel<<<blocks, kFillBlock, 0, s>>>(idx, mass, pos, n, t);
    if (n > kFillBlock) {
        const uint32_t num_super = num_supers(n);
        fill_mid_kernel<<<num_super, kFillBlock, 0, s>>>(n, t, num_super,
                                                         climb ? 0u : kLdsList);
    }
}

int64_t fill_count_words(int64_t cap) { return 2 + (int64_t)num_supers(cap); }
int64_t fill_list_words(int64_t cap) { return (int64_t)num_supers(cap) * kTopPer; }

}  // namespace nbs
