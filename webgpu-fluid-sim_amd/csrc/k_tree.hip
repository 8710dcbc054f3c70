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
//     phase 3 (one more fill_mid workgroup, at the same time): completes the
//       top nodes, each climb waiting only for the super-blocks whose
//       phase-2 records it merges.
//   A kernel boundary orders phases 1 and 2; a device-scope fence and a
//   done flag per phase-2 workgroup order phases 2 and 3.
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
constexpr int kBuildThreads = 512;
constexpr int kBuildPerThread = 1;  // nodes per thread (2: all wavefronts resident at once, measured slower)
constexpr int kBuildNodes = kBuildThreads * kBuildPerThread;  // nodes per workgroup
constexpr int kBuildHalo = 512;
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
    // a position known to lie in the window
    __device__ __forceinline__ K lds(int j) const { return s[j - lo]; }
};
// Load the window of the workgroup's nodes [b0, b0 + kBuildNodes) into s_keys.
template <typename K>
__device__ __forceinline__ KeyWindow<K> load_window(const K *__restrict__ keys, int n,
                                                    K *s_keys) {
    const int b0 = (int)blockIdx.x * kBuildNodes;
    const int lo = b0 > kBuildHalo ? b0 - kBuildHalo : 0;
    const int hi = b0 + kBuildNodes + kBuildHalo < n ? b0 + kBuildNodes + kBuildHalo : n;
    // every load issued before the first LDS write: one memory round trip
    constexpr int kPer = (kBuildWindow + kBuildThreads - 1) / kBuildThreads;
    K v[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int j = lo + (int)threadIdx.x + r * kBuildThreads;
        v[r] = j < hi ? keys[j] : K(0);
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int j = lo + (int)threadIdx.x + r * kBuildThreads;
        if (j < hi) s_keys[j - lo] = v[r];
    }
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
        __device__ __forceinline__ int cmp(int j, uint32_t kj) const {
            const uint32_t x = ki ^ kj;
            const int by_key = __clz((int)x), by_pos = 32 + __clz(i ^ j);
            return x ? by_key : by_pos;
        }
        __device__ __forceinline__ int operator()(int j) const {
            if ((uint32_t)j >= (uint32_t)d.n) return -1;
            return cmp(j, d.k[j]);
        }
        // j in the window (or outside [0, n))
        __device__ __forceinline__ int local(int j) const {
            if ((uint32_t)j >= (uint32_t)d.n) return -1;
            return cmp(j, d.k.lds(j));
        }
    };
    // i in the window
    __device__ __forceinline__ Node at(int i) const { return Node{*this, k.lds(i), i}; }
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
        __device__ __forceinline__ int cmp(int j, uint64_t b) const {
            if (ki == b) {
                const int off = (int)d.sim_off[ki >> kSimShift];
                return 64 + __clz((i - off) ^ (j - off));
            }
            return __clzll((long long)(ki ^ b));
        }
        __device__ __forceinline__ int operator()(int j) const {
            if ((uint32_t)j >= (uint32_t)d.n) return -1;
            return cmp(j, d.k[j]);
        }
        // j in the window (or outside [0, n))
        __device__ __forceinline__ int local(int j) const {
            if ((uint32_t)j >= (uint32_t)d.n) return -1;
            return cmp(j, d.k.lds(j));
        }
    };
    // i in the window
    __device__ __forceinline__ Node at(int i) const { return Node{*this, k.lds(i), i}; }
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
// The fill's upper levels: the build lists every node whose leaf range
// spans a block boundary but lies in one super-block in that super-block's
// phase-2 list, and every node whose range spans super-blocks in the top
// list (fill_mid_kernel resolves them); tb.slot holds each listed node's
// position in its list.  A node's index is an end of its range, so a
// phase-2 node of super-block s has its index in [s * kSuper, (s+1) *
// kSuper): a phase-2 list has at most kSuper entries.  The top nodes whose
// index lies in one super-block are the nodes whose range holds its left
// or right boundary, two root-ward chains of at most the tree depth (< 97
// for 38-bit keys tie-broken on positions below 2^30), so the top list has
// fewer than kTopPer entries per super-block.
constexpr uint32_t kLdsList = 1024;  // longest list the fill resolves in LDS
__host__ __device__ inline uint32_t num_supers(int64_t n) {
    return (uint32_t)((n + kSuper - 1) / kSuper);
}
constexpr uint32_t kTopPer = 256;
// tb.counts: the length of list l (phase-2 list of super-block l < S, the
// top list l = S) at list_count(l), one 128-byte line each, as the builds'
// atomics on one line would serialize; the next word of super-block l's
// line is its done flag (its phase-2 nodes are published).
// tb.lists: the phase-2 lists (kSuper words each), then the top list.
__host__ __device__ inline uint32_t list_count(uint32_t l) { return 32u * (l + 1u); }
__host__ __device__ inline uint32_t done_flag(uint32_t sb) { return list_count(sb) + 1u; }

__device__ __forceinline__ bool spans(uint2 r, int shift) { return (r.x >> shift) != (r.y >> shift); }

// kInBlock when the node's leaf range lies in one phase-1 block.
__device__ __forceinline__ uint32_t in_block(const Split &sp) {
    return ((uint32_t)sp.lo >> kBlockShift) == ((uint32_t)sp.hi >> kBlockShift) ? kInBlock : 0u;
}

// Add node i (leaf range sp.lo..sp.hi) to its upper-level list, if any.
// Called by every active lane of the wavefront; the lanes adding to one
// list share one atomic.  A list is in no particular order.
__device__ __forceinline__ void list_node(const TreeBufs &tb, uint32_t i, const Split &sp,
                                          int n) {
    const uint32_t lo = (uint32_t)sp.lo, hi = (uint32_t)sp.hi;
    const bool top = (lo >> kSuperShift) != (hi >> kSuperShift);
    const bool listed = top || (lo >> kBlockShift) != (hi >> kBlockShift);
    const uint32_t S = num_supers(n);
    const uint32_t lid = top ? S : (i >> kSuperShift);  // S: the top list
    const int lane = threadIdx.x & 63;
    uint64_t todo = __ballot(listed);
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)lid, leader);
        const uint64_t peers = __ballot(listed && lid == l);
        uint32_t pos0 = 0;
        if (lane == leader)
            pos0 = atomicAdd(&tb.counts[list_count(l)], (uint32_t)__builtin_popcountll(peers));
        pos0 = (uint32_t)__builtin_amdgcn_readlane((int)pos0, leader);
        if (listed && lid == l) {
            const uint32_t pos =
                pos0 + __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
            tb.lists[(size_t)l * kSuper + pos] = i;  // the top list follows the S phase-2 lists
            tb.slot[i] = pos;
        }
        todo &= ~peers;
    }
}

__device__ __forceinline__ int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
// a key from a per-lane source lane
__device__ __forceinline__ uint32_t shfl_key(uint32_t v, int src) {
    return (uint32_t)__shfl((int)v, src);
}
__device__ __forceinline__ uint64_t shfl_key(uint64_t v, int src) {
    return (uint64_t)(uint32_t)__shfl((int)(uint32_t)v, src) |
           ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(v >> 32), src) << 32);
}
__device__ __forceinline__ uint32_t bcast(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ uint64_t bcast(uint64_t v, int lane) {
    return (uint64_t)bcast((uint32_t)v, lane) | ((uint64_t)bcast((uint32_t)(v >> 32), lane) << 32);
}

// The last step x in [lo, hi) along direction d from position i at which the
// common prefix with key ki exceeds `bound`, given that it does at lo and
// does not at hi (every argument wavefront-uniform, every lane active).
// Each round the 64 lanes probe 63 evenly spaced steps at once, so R
// positions take log64(R) rounds of independent loads where a binary
// search takes log2(R) dependent ones.
template <typename Delta, typename K>
__device__ __forceinline__ int wave_search(const Delta &delta, K ki, int i, int d, int bound, int lo,
                                           int hi) {
    const int lane = threadIdx.x & 63;
    const typename Delta::Node nd{delta, ki, i};
    while (hi - lo > 1) {
        const int step = (hi - lo + 63) / 64;
        const int x = lo + lane * step;
        const bool p = lane > 0 && x < hi && nd(d > 0 ? i + x : i - x) > bound;
        // the prefix only shrinks with distance: lanes 1..c hold
        const int c = __builtin_popcountll(__ballot(p));
        const int next_lo = lo + c * step;
        hi = min(lo + (c + 1) * step, hi);
        lo = next_lo;
    }
    return lo;
}

// Two such searches at once, one per half-wavefront of 32 lanes (31
// probes per round each): the two wide nodes a wavefront most often holds
// (the children at one split, of neighbouring indices) take the rounds of
// one search instead of two.  Arguments per lane: those of the search its
// half serves (uniform within each half); every lane active.  Returns each
// half's result (uniform within the half).
template <typename Delta, typename K>
__device__ __forceinline__ int wave_search2(const Delta &delta, K ki, int i, int d, int bound,
                                            int lo, int hi) {
    const int lane = threadIdx.x & 63, h = lane & 31;
    const typename Delta::Node nd{delta, ki, i};
    while (__ballot(hi - lo > 1)) {
        const bool open = hi - lo > 1;
        const int step = (hi - lo + 31) / 32;
        const int x = lo + h * step;
        const bool p = open && h > 0 && x < hi && nd(d > 0 ? i + x : i - x) > bound;
        const uint64_t b = __ballot(p);
        const int c = __builtin_popcount((uint32_t)(lane < 32 ? b : b >> 32));
        if (open) {
            const int next_lo = lo + c * step;
            hi = min(lo + (c + 1) * step, hi);
            lo = next_lo;
        }
    }
    return lo;
}

// lbvh.wgsl:101-152 for node i (valid lanes; every lane of the wavefront
// calls it).  Both searches look for the end of a run of steps whose common
// prefix with key i exceeds a bound (the run ends at a unique step, so any
// correct search returns the reference's split).  A lane searches inside the
// workgroup's key window (LDS); the nodes whose answer lies beyond it (wide
// leaf ranges: about one node in 256) are searched by the whole wavefront
// together, one node after the other.  Probe i + d * x is written as a
// select on d.
template <typename Delta>
__device__ __forceinline__ Split karras_node(const Delta &delta, int i, int n, bool valid) {
    using Node = typename Delta::Node;
    using K = decltype(Node::ki);
    const int lane = threadIdx.x & 63;
    K ki = 0;
    int d = 1, delta_min = 0, xw = 0, l = 0;
    bool far = false;
    if (valid) {
        const Node dl = delta.at(i);
        ki = dl.ki;
        // every probe of this part lies in the window
        d = dl.local(i + 1) > dl.local(i - 1) ? 1 : -1;
        delta_min = dl.local(i - d);
        // the farthest step inside the window, and inside the array
        xw = d > 0 ? delta.k.lo + delta.k.len - 1 - i : i - delta.k.lo;
        const int xa = d > 0 ? n - 1 - i : i;
        int l_max = 2;
        while (l_max <= xw && dl.local(d > 0 ? i + l_max : i - l_max) > delta_min) l_max *= 2;
        far = l_max > xw && xw < xa && dl.local(d > 0 ? i + xw : i - xw) > delta_min;
        if (!far)  // the range ends inside the window: steps beyond xw fail
            for (int t = l_max >> 1; t >= 1; t >>= 1)
                if (l + t <= xw && dl.local(d > 0 ? i + l + t : i - l - t) > delta_min) l += t;
    }
    uint64_t fm = __ballot(far);
    if (__builtin_popcountll(fm) == 2) {
        // both at once, the first one's in lanes 0-31, the second one's in 32-63
        const int l0 = __builtin_ctzll(fm), l1 = __builtin_ctzll(fm & (fm - 1));
        const int ld = lane < 32 ? l0 : l1;
        const int di = __shfl(d, ld), ii = __shfl(i, ld);
        const int r = wave_search2(delta, shfl_key(ki, ld), ii, di, __shfl(delta_min, ld),
                                   __shfl(xw, ld), (di > 0 ? n - 1 - ii : ii) + 1);
        const int r0 = bcast(r, 0), r1 = bcast(r, 32);
        if (lane == l0) l = r0;
        if (lane == l1) l = r1;
        fm = 0;
    }
    for (uint64_t m = fm; m; m &= m - 1) {
        const int ld = __builtin_ctzll(m);
        const int xa = bcast(d, ld) > 0 ? n - 1 - bcast(i, ld) : bcast(i, ld);
        const int r = wave_search(delta, bcast(ki, ld), bcast(i, ld), bcast(d, ld),
                                  bcast(delta_min, ld), bcast(xw, ld), xa + 1);
        if (lane == ld) l = r;
    }
    int s = 0, delta_node = 0;
    bool far2 = false;
    if (valid) {
        const Node dl{delta, ki, i};
        delta_node = dl(d > 0 ? i + l : i - l);
        int bound = l;
        if (l > xw) {  // a wide node: is the split beyond the window too?
            far2 = dl.local(d > 0 ? i + xw : i - xw) > delta_node;
            bound = xw;
        }
        if (!far2) {
            // the test fails at every step >= bound (the range's end, or a
            // window end past which the prefix is no longer than delta_node),
            // so steps beyond it are not probed and every probe lies in the
            // window
            int t = (bound + 1) >> 1;
            for (;;) {
                if (s + t <= bound && dl.local(d > 0 ? i + s + t : i - s - t) > delta_node) s += t;
                if (t == 1) break;
                t = (t + 1) >> 1;
            }
        }
    }
    uint64_t fm2 = __ballot(far2);
    if (__builtin_popcountll(fm2) == 2) {
        const int l0 = __builtin_ctzll(fm2), l1 = __builtin_ctzll(fm2 & (fm2 - 1));
        const int ld = lane < 32 ? l0 : l1;
        const int r = wave_search2(delta, shfl_key(ki, ld), __shfl(i, ld), __shfl(d, ld),
                                   __shfl(delta_node, ld), __shfl(xw, ld), __shfl(l, ld));
        const int r0 = bcast(r, 0), r1 = bcast(r, 32);
        if (lane == l0) s = r0;
        if (lane == l1) s = r1;
        fm2 = 0;
    }
    for (uint64_t m = fm2; m; m &= m - 1) {
        const int ld = __builtin_ctzll(m);
        const int r = wave_search(delta, bcast(ki, ld), bcast(i, ld), bcast(d, ld),
                                  bcast(delta_node, ld), bcast(xw, ld), bcast(l, ld));
        if (lane == ld) s = r;
    }
    Split r{};
    const int j = d > 0 ? i + l : i - l;
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
// (kThreads: the workgroup size, a constant, so that no wave waits for the
// dispatch packet's copy of it)
template <int kThreads>
__device__ __forceinline__ void clear_words(uint32_t *__restrict__ p, int64_t words,
                                            uint32_t *__restrict__ slots) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * kThreads)
        p[i] = 0u;
    if (slots && blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < (uint32_t)kBBoxSlots; k += kThreads)
            reinterpret_cast<uint4 *>(slots)[k] = make_uint4(0u, 0u, 0u, 0u);
}

// Leaf record {x, y, mass, 0} of sorted position i (lbvh.wgsl:179-187),
// gathered by body index: the fill's way for a single body (no sort, no
// build).  Otherwise the build places every body's record at its sorted
// position (place_leaf).
__device__ __forceinline__ float4 gather_leaf(const TreeBufs &tb, const uint32_t *__restrict__ idx,
                                              const float *__restrict__ mass,
                                              const float2 *__restrict__ pos, int i) {
    const uint32_t b = idx[i];
    const float2 p = pos[b];
    const float4 c = make_float4(p.x, p.y, mass[b], 0.0f);
    tb.leaf[i] = c;
    return c;
}

// Body b's leaf record at its sorted position: coalesced reads by body, one
// scattered 16-byte write.
__device__ __forceinline__ void place_leaf(const TreeBufs &tb, const LeafPlace &lpl, int b) {
    const float2 p = lpl.pos[b];
    tb.leaf[lpl.rank[b]] = make_float4(p.x, p.y, lpl.mass[b], 0.0f);
}

__global__ __launch_bounds__(kBuildThreads) void build_u32_kernel(
    const uint32_t *__restrict__ keys, int64_t n, TreeBufs tb, LeafPlace lpl,
    uint32_t *__restrict__ clear, int64_t clear_n, uint32_t *__restrict__ slots) {
    __shared__ uint32_t s_keys[kBuildWindow];
    clear_words<kBuildThreads>(clear, clear_n, slots);
    const int n32 = (int)n;
    const DeltaU32 delta{load_window(keys, n32, s_keys), n32};
    for (int r = 0; r < kBuildPerThread; ++r) {
        const int i = (int)blockIdx.x * kBuildNodes + r * kBuildThreads + (int)threadIdx.x;
        const bool valid = i < n32 - 1;
        // body i's leaf record: read before the searches, written after them
        // (the searches wait on LDS only, so the reads' latency is hidden)
        float4 leaf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t at = 0u;
        if (valid) {
            const float2 p = lpl.pos[i];
            leaf = make_float4(p.x, p.y, lpl.mass[i], 0.0f);
            at = lpl.rank[i];
        }
        // every lane takes part in its wavefront's searches of wide nodes
        const Split sp = karras_node(delta, i, n32, valid);
        if (!valid) return;
        tb.leaf[at] = leaf;
        if (i == n32 - 2) place_leaf(tb, lpl, n32 - 1);
        if (lpl.split) {  // the splitters of the next step's bucket sort
            const int every = lpl.split_every;
            if (i > 0 && i % every == 0) lpl.split[i / every - 1] = keys[i];
            if (i == n32 - 2 && (n32 - 1) % every == 0)
                lpl.split[(n32 - 1) / every - 1] = keys[n32 - 1];
        }
        list_node(tb, (uint32_t)i, sp, n32);
        tb.child[i] = make_uint2(sp.lc, sp.rc);
        // the walk reads the indices from the pair record; the fill writes
        // the records of nodes inside one block whole
        const uint32_t ib = in_block(sp);
        if (!ib)
            *reinterpret_cast<uint2 *>(tb.pcm + kPairStride * (size_t)i + 2) =
                make_uint2(sp.lc, sp.rc);
        tb.range[i] = make_uint2((uint32_t)sp.lo, (uint32_t)sp.hi);
        // lbvh.wgsl:155-161, with the fill's in-block flag
        const uint32_t pw = (uint32_t)i | ib;
        tb.parent[sp.lc] = pw;
        tb.parent[sp.rc] = pw | kRightChild;
        if (i == 0) tb.parent[0] = kNoNode;
        tb.status[i] = 0u;  // lbvh.wgsl:164
    }
}

__global__ __launch_bounds__(kBuildThreads) void build_u64_kernel(const uint64_t *__restrict__ keys,
                                                        int64_t n,
                                                        const int64_t *__restrict__ sim_off,
                                                        TreeBufs tb, LeafPlace lpl,
                                                        uint32_t *__restrict__ sim_root,
                                                        uint32_t *__restrict__ clear,
                                                        int64_t clear_n,
                                                        uint32_t *__restrict__ slots) {
    __shared__ uint64_t s_keys[kBuildWindow];
    clear_words<kBuildThreads>(clear, clear_n, slots);
    const int n32 = (int)n;
    const DeltaU64 delta{load_window(keys, n32, s_keys), sim_off, n32};
    for (int r = 0; r < kBuildPerThread; ++r) {
        const int i = (int)blockIdx.x * kBuildNodes + r * kBuildThreads + (int)threadIdx.x;
        const bool valid = i < n32 - 1;
        // body i's leaf record: read before the searches, written after them
        // (the searches wait on LDS only, so the reads' latency is hidden)
        float4 leaf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t at = 0u;
        if (valid) {
            const float2 p = lpl.pos[i];
            leaf = make_float4(p.x, p.y, lpl.mass[i], 0.0f);
            at = lpl.rank[i];
        }
        // every lane takes part in its wavefront's searches of wide nodes
        const Split sp = karras_node(delta, i, n32, valid);
        if (!valid) return;
        tb.leaf[at] = leaf;
        if (i == n32 - 2) place_leaf(tb, lpl, n32 - 1);
        list_node(tb, (uint32_t)i, sp, n32);
        tb.child[i] = make_uint2(sp.lc, sp.rc);
        const uint32_t ib = in_block(sp);
        if (!ib)  // (the fill writes in-block records whole)
            *reinterpret_cast<uint2 *>(tb.pcm + kPairStride * (size_t)i + 2) =
                make_uint2(sp.lc, sp.rc);
        tb.range[i] = make_uint2((uint32_t)sp.lo, (uint32_t)sp.hi);
        const uint32_t pw = (uint32_t)i | ib;
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
    if (tb.pair_keys)
        reinterpret_cast<float *>(tb.pcm)[16 * (size_t)(pw & kParentMask) + 10 + (pw >> 31)] =
            open_key(C.w, tb.open_scale, tb.open_bias);
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
    __shared__ uint32_t s_arrive[kFillBlock];  // arrivals; 2 once the node is complete
    __shared__ uint32_t s_start[kFillBlock];   // the climbs' start nodes, compacted
    __shared__ uint32_t s_nstart;

    const uint32_t j = threadIdx.x;
    const uint32_t base = blockIdx.x * kFillBlock;
    const uint32_t nint = (uint32_t)(n - 1);  // internal nodes
    const uint32_t t = base + j;
    s_arrive[j] = 0u;
    if (j == 0) s_nstart = 0u;
    // every load before the first LDS write of a loaded value: one memory
    // round trip
    // (the leaf's first: the compiler moves the start test below up to the
    // node's range, and its wait then comes after every load is issued)
    uint32_t pw_leaf = kNoNode;
    float4 leaf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (t < (uint64_t)n) {
        pw_leaf = n > 1 ? tb.parent[nint + t] : kNoNode;
        // the build placed the leaf records; a single body has no build
        leaf = n == 1 ? gather_leaf(tb, idx, mass, pos, 0) : tb.leaf[t];
    }
    uint2 ch = make_uint2(0u, 0u), rg = make_uint2(0u, 0u);
    uint32_t par_int = 0u;
    if (t < nint) {
        ch = tb.child[t];
        rg = tb.range[t];
        par_int = tb.parent[t];
    }
    if (t < nint) {
        s_child[j] = ch;
        s_par_int[j] = par_int;
    }
    if (t < (uint64_t)n) s_cm[kFillBlock + j] = leaf;
    __syncthreads();

    // lbvh.wgsl:191-232, restricted to the nodes whose leaf range lies in
    // this block.  A climb starts at each such node with two leaf children
    // and carries the node in registers.  At the parent, a leaf sibling's
    // record is in LDS from the start; an internal sibling is another
    // climb's, so the two meet at an arrival count and the second one
    // climbs on.  A level is one read of the parent's child pair and
    // parent word, then of the sibling (after the count for an internal
    // sibling), and the merge; a parent word without kInBlock ends it.  The
    // start nodes are gathered into the first threads, so only the few
    // wavefronts that climb are active.
    if (t < nint && ch.x >= nint && ch.y >= nint && !spans(rg, kBlockShift))
        s_start[atomicAdd(&s_nstart, 1u)] = j;
    __syncthreads();
    if (j < s_nstart) {
        const uint32_t c0 = s_start[j];
        const uint2 cch = s_child[c0];
        const float4 L = s_cm[kFillBlock + (cch.x - nint - base)];
        const float4 R = s_cm[kFillBlock + (cch.y - nint - base)];
        float4 C, B;
        merge_values(L, make_float4(L.x, L.y, L.x, L.y), R, make_float4(R.x, R.y, R.x, R.y), C, B);
        uint32_t cur = c0, pw = s_par_int[c0];
        for (int level = 0; level < 128; ++level) {
            s_cm[cur] = C;
            s_bb[cur] = B;
            s_arrive[cur] = 2u;  // complete (no climb arrives here any more)
            if (base + cur == 0u || !(pw & kInBlock)) break;  // the root, or a block top
            const uint32_t lp = (pw & kParentMask) - base;
            if (lp >= (uint32_t)kFillBlock) break;  // corrupt parent word
            const uint2 pch = s_child[lp];
            const uint32_t ppw = s_par_int[lp];
            const bool right = (pw >> 31) != 0u;
            const uint32_t sib = right ? pch.x : pch.y;
            float4 S, SB;
            if (sib < nint) {
                if (__hip_atomic_fetch_add(&s_arrive[lp], 1u, __ATOMIC_ACQ_REL,
                                           __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
                    break;  // the sibling's climb is still on its way: it goes on
                S = s_cm[sib - base];
                SB = s_bb[sib - base];
            } else {
                S = s_cm[kFillBlock + (sib - nint - base)];
                SB = make_float4(S.x, S.y, S.x, S.y);
            }
            float4 NC, NB;
            merge_values(right ? S : C, right ? SB : B, right ? C : S, right ? B : SB, NC, NB);
            C = NC;
            B = NB;
            cur = lp;
            pw = ppw;
        }
    }
    __syncthreads();

    // Write the block's completed nodes where their readers look for them.
    // A completed internal node writes its own pair record whole (its
    // children are complete in this block): one 64-byte record and one
    // 32-byte AABB pair per node, no partial lines.  Only the nodes whose
    // parent lies outside the block (block tops, and leaves or nodes under a
    // spanning parent) publish themselves into their parent's slots, which
    // the upper phases read.
    if (t < (uint64_t)n) {
        float4 c = s_cm[kFillBlock + j];
        c.w = kLeafLength;
        if (n == 1)
            publish(tb, 0u, 0u, c, make_float4(c.x, c.y, c.x, c.y));
        else if (!(pw_leaf & kInBlock))
            publish(tb, nint + t, pw_leaf, c, make_float4(c.x, c.y, c.x, c.y));
    }
    if (t < nint && s_arrive[j] == 2u) {
        const uint32_t pw = t ? s_par_int[j] : 0u;
        if (t == 0u || !(pw & kInBlock)) publish(tb, t, pw, s_cm[j], s_bb[j]);
        float4 cd[2], bd[2];
        const uint32_t cc[2] = {ch.x, ch.y};
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            if (cc[side] >= nint) {
                cd[side] = s_cm[kFillBlock + (cc[side] - nint - base)];
                cd[side].w = kLeafLength;
                bd[side] = make_float4(cd[side].x, cd[side].y, cd[side].x, cd[side].y);
            } else {
                cd[side] = s_cm[cc[side] - base];
                bd[side] = s_bb[cc[side] - base];
            }
        }
        float4 *rec = tb.pcm + kPairStride * (size_t)t;
        rec[0] = cd[0];
        rec[1] = cd[1];
        float4 w2 = make_float4(__uint_as_float(ch.x), __uint_as_float(ch.y), 0.0f, 0.0f);
        if (tb.pair_keys) {
            w2.z = open_key(cd[0].w, tb.open_scale, tb.open_bias);
            w2.w = open_key(cd[1].w, tb.open_scale, tb.open_bias);
        }
        rec[2] = w2;
        rec[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        tb.paabb[2 * (size_t)t] = bd[0];
        tb.paabb[2 * (size_t)t + 1] = bd[1];
    }
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

// One upper-level node list in LDS: entry e is node node[e] with parent
// word pw[e], its parent's entry par[e] (kNone: the parent is not in the
// list), the number of its children that are entries (pend), an arrival
// count, and its children's {cm, aabb} records (left, right).  src[e]
// (top list): per side (bits 0-15 left, 16-31 right), 1 + the super-block
// whose phase-2 workgroup produces that child's record, or 0.
constexpr uint32_t kNone = 0xFFFFFFFFu;
struct ListLds {
    uint32_t node[kLdsList], pw[kLdsList], par[kLdsList], pend[kLdsList], cnt[kLdsList];
    uint32_t src[kLdsList];
    uint32_t start[kLdsList], nstart;  // the entries with no child among the entries
    float4 in[kLdsList][4];
    float4 out[kLdsList][2];  // the entry's own {cm, aabb} once complete
};

// A float4 read at device scope (past this XCD's L2, which may hold the
// line from before another XCD's write).
__device__ __forceinline__ float4 load_device_f4(const float4 *p) {
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
    const uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)),
                       __uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32)));
}

// The record of child `side` of top node `node` when a phase-2 workgroup
// produces it: wait for that super-block's done flag (sb1 = 1 + the
// super-block), then read the record at device scope.
__device__ __forceinline__ void wait_input(const TreeBufs &tb, uint32_t sb1, uint32_t node,
                                           uint32_t side, float4 &V, float4 &VB) {
    const uint32_t *flag = tb.counts + done_flag(sb1 - 1u);
    for (uint32_t spins = 0;
         __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u; ++spins) {
        if (spins == kSpinLimit) {
            atomicOr(tb.fault, kFaultWait);
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    // acquire (agent scope): pairs with the producer's release fence before
    // its flag store, so the records below are the producer's
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    V = load_device_f4(tb.pcm + kPairStride * (size_t)node + side);
    VB = load_device_f4(tb.paabb + 2 * (size_t)node + side);
}

// Resolve a list set up in ls (m <= kLdsList entries; called by the whole
// workgroup).  The entries with no child among the entries start a climb in
// LDS that carries the node in registers: merge the children (lbvh.wgsl:
// 205-228), keep the node in its entry and go to the parent entry.  A
// parent with one entry child has its other input preloaded, so the climb
// goes straight on; a parent with two meets both climbs at an arrival count
// (workgroup-scope acquire/release), and the second one goes on.  The
// chains of the upper levels are long (tens of levels, mostly parents with
// one entry child), so the nodes are published to memory only afterwards,
// by all threads together.  kTop: the top list, whose inputs from the
// phase-2 workgroups (src) are read when their super-block is done.
template <bool kTop>
__device__ __forceinline__ void resolve_list(const TreeBufs &tb, ListLds &ls, uint32_t m) {
    // The climbs start from the entries without entry children, gathered into
    // the first threads: the few wavefronts that climb (one for the top list)
    // then share their CU with no other climbing wavefront.
    if (threadIdx.x == 0) ls.nstart = 0u;
    __syncthreads();
    if (threadIdx.x < m && ls.pend[threadIdx.x] == 0u)
        ls.start[atomicAdd(&ls.nstart, 1u)] = threadIdx.x;
    __syncthreads();
    const uint32_t nstart = ls.nstart;
    // Top list: the last wavefront loads the phase-2 inputs into LDS as their
    // super-blocks finish and then clears that side of the entry's src
    // (release), so a climb waits on LDS; when the climbs need that wavefront
    // too, each climb reads its own phase-2 inputs.  A wavefront's LDS
    // operations execute in order, so a climb that reads a side of src as 0
    // and then that input reads it complete.  Each loader lane polls the done
    // flags of all its pending inputs (lane l: entries l + 64k, bit 2k + side)
    // and loads every input whose super-block is done, behind one acquire per
    // round: the last super-block's inputs are one round away from its flag,
    // whichever lanes hold them.
    constexpr uint32_t kLoader = kFillBlock - 64;
    const bool loaders = kTop && nstart <= kLoader;
    if (kTop && loaders && threadIdx.x >= kLoader) {
        const uint32_t l = threadIdx.x - kLoader;  // m <= kLdsList: k < 16
        uint32_t pend = 0u;
        for (uint32_t k = 0; l + 64u * k < m; ++k) {
            const uint32_t sc = ls.src[l + 64u * k];
            pend |= ((sc & 0xFFFFu) ? 1u : 0u) << (2u * k);
            pend |= ((sc >> 16) ? 1u : 0u) << (2u * k + 1u);
        }
        for (uint32_t spins = 0; __ballot(pend != 0u); ++spins) {
            const bool give_up = spins == kSpinLimit;
            uint32_t ready = 0u;
            for (uint32_t b = pend; b; b &= b - 1u) {
                const uint32_t bit = (uint32_t)__builtin_ctz(b), e = l + 64u * (bit >> 1);
                const uint32_t sb1 = (ls.src[e] >> (16u * (bit & 1u))) & 0xFFFFu;
                if (give_up || __hip_atomic_load(tb.counts + done_flag(sb1 - 1u), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) != 0u)
                    ready |= 1u << bit;
            }
            if (give_up && pend) atomicOr(tb.fault, kFaultWait);
            if (!__ballot(ready != 0u)) {
                __builtin_amdgcn_s_sleep(2);
                continue;
            }
            // acquire (agent scope): pairs with the producers' release fences
            // before their flag stores
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            for (uint32_t b = ready; b; b &= b - 1u) {
                const uint32_t bit = (uint32_t)__builtin_ctz(b), e = l + 64u * (bit >> 1);
                const uint32_t side = bit & 1u, node = ls.node[e];
                ls.in[e][2 * side] = load_device_f4(tb.pcm + kPairStride * (size_t)node + side);
                ls.in[e][2 * side + 1] = load_device_f4(tb.paabb + 2 * (size_t)node + side);
                __hip_atomic_fetch_and(&ls.src[e], ~(0xFFFFu << (16u * side)), __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            pend &= ~ready;
        }
    }
    // input `side` of entry e, produced by a phase-2 workgroup
    auto mid_input = [&](uint32_t e, uint32_t side, float4 &V, float4 &VB) {
        if (loaders) {
            while ((__hip_atomic_load(&ls.src[e], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >>
                    (16u * side)) &
                   0xFFFFu)
                __builtin_amdgcn_s_sleep(1);
            V = ls.in[e][2 * side];
            VB = ls.in[e][2 * side + 1];
        } else {
            wait_input(tb, (ls.src[e] >> (16u * side)) & 0xFFFFu, ls.node[e], side, V, VB);
        }
    };
    if (threadIdx.x < nstart) {
        uint32_t cur = ls.start[threadIdx.x];
        uint32_t sc = 0u;
        if constexpr (kTop) {
            sc = __hip_atomic_load(&ls.src[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            asm volatile("" ::: "memory");  // src before the inputs
        }
        float4 L = ls.in[cur][0], LB = ls.in[cur][1], R = ls.in[cur][2], RB = ls.in[cur][3];
        uint32_t pw = ls.pw[cur], q = ls.par[cur];
        if constexpr (kTop) {
            if (sc & 0xFFFFu) mid_input(cur, 0u, L, LB);
            if (sc >> 16) mid_input(cur, 1u, R, RB);
        }
        for (int level = 0; level < 128; ++level) {
            // the parent's fixed words and its other input are read before
            // the merge, which hides their latency; they are read
            // unconditionally (entry 0 stands in above the list's last
            // entry, and is not used), so no test waits for them
            const uint32_t side = pw >> 31;
            const uint32_t qs = q == kNone ? 0u : q;
            uint32_t sq = 0u;
            if constexpr (kTop) {
                sq = (__hip_atomic_load(&ls.src[qs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >>
                      (16u * (side ^ 1u))) &
                     0xFFFFu;
                asm volatile("" ::: "memory");  // src before the input
            }
            const uint32_t qpw = ls.pw[qs], qq = ls.par[qs], qpend = ls.pend[qs];
            float4 S = ls.in[qs][2 * (side ^ 1u)];
            float4 SB = ls.in[qs][2 * (side ^ 1u) + 1];
            float4 C, B;
            merge_values(L, LB, R, RB, C, B);
            ls.out[cur][0] = C;
            ls.out[cur][1] = B;
            if (q == kNone) break;
            if constexpr (kTop) {
                if (sq) mid_input(q, side ^ 1u, S, SB);
            }
            if (qpend == 2u) {
                // both children are entries: the one that arrives second
                // climbs on, with the other's record read after the count
                ls.in[q][2 * side] = C;
                ls.in[q][2 * side + 1] = B;
                if (__hip_atomic_fetch_add(&ls.cnt[q], 1u, __ATOMIC_ACQ_REL,
                                           __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
                    break;
                S = ls.in[q][2 * (side ^ 1u)];
                SB = ls.in[q][2 * (side ^ 1u) + 1];
            }
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

// Set up entry t < m of a list of nodes (list[t]) in ls: its parent's
// entry (tb.slot) when the parent is listed, the number of its children
// that are entries, and both children's records as the last fill phase
// left them (an entry child's is replaced by its climb).  `shift`:
// kBlockShift for a phase-2 list, kSuperShift for the top list.
__device__ __forceinline__ void setup_entry(const TreeBufs &tb, ListLds &ls,
                                            const uint32_t *list, uint32_t t, uint32_t nint,
                                            int shift) {
    const uint32_t p = list[t];
    const uint32_t pw = tb.parent[p];
    const uint2 ch = tb.child[p];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        ls.in[t][2 * side] = tb.pcm[kPairStride * (size_t)p + side];
        ls.in[t][2 * side + 1] = tb.paabb[2 * (size_t)p + side];
    }
    // the parent's range and slot and the children's ranges in one round
    // trip (an index outside the internal nodes reads node 0, unused)
    const uint32_t par = pw & kParentMask;
    const bool par_in = par < nint;
    const uint2 rpar = tb.range[par_in ? par : 0u];
    const uint32_t spar = tb.slot[par_in ? par : 0u];
    const uint2 rch[2] = {tb.range[ch.x < nint ? ch.x : 0u], tb.range[ch.y < nint ? ch.y : 0u]};
    uint32_t q = kNone;
    // a top node's parent is a top node; a phase-2 node's parent is listed
    // with it unless it spans super-blocks
    if (par_in && (shift == kSuperShift || !spans(rpar, kSuperShift))) q = spar;
    uint32_t pend = 0u, src = 0u;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const uint32_t c = side ? ch.y : ch.x;
        if (c >= nint) continue;
        const uint2 rc = rch[side];
        if (spans(rc, shift))
            ++pend;
        else if (shift == kSuperShift && spans(rc, kBlockShift))
            src |= ((rc.x >> kSuperShift) + 1u) << (16 * side);  // a phase-2 node
    }
    ls.node[t] = p;
    ls.pw[t] = pw;
    ls.par[t] = q;
    ls.pend[t] = pend;
    ls.src[t] = src;
    ls.cnt[t] = 0u;
}

// A top list longer than kLdsList is resolved in parts once every phase-2
// workgroup is done: part h < H holds the top nodes whose leaf range lies in
// hyper-block h (2^kHyperShift leaves), part H the ones spanning
// hyper-blocks.  Set up entry t of a part (cross: part H); tb.slot holds
// each node's entry in its part.  Every input that is not an entry of the
// part is in memory: phase 1's records and the earlier parts' as written by
// this workgroup, the phase-2 workgroups' at device scope.
constexpr int kHyperShift = 20;
__device__ __forceinline__ void setup_part_entry(const TreeBufs &tb, ListLds &ls,
                                                 const uint32_t *part, uint32_t t, uint32_t nint,
                                                 bool cross) {
    const uint32_t p = part[t];
    const uint32_t pw = tb.parent[p];
    const uint2 ch = tb.child[p];
    const uint32_t par = pw & kParentMask;
    // the parent's range and slot and the children's ranges in one round
    // trip (an index outside the internal nodes reads node 0, unused)
    const bool par_in = par < nint;
    const uint2 rpar = tb.range[par_in ? par : 0u];
    const uint32_t spar = tb.slot[par_in ? par : 0u];
    const uint2 rch[2] = {tb.range[ch.x < nint ? ch.x : 0u], tb.range[ch.y < nint ? ch.y : 0u]};
    uint32_t q = kNone;
    // the parent, a top node, is in the part when it lies in the same
    // hyper-block as the node, or when both span hyper-blocks
    if (par_in && (cross || !spans(rpar, kHyperShift))) q = spar;
    uint32_t pend = 0u;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const uint32_t c = side ? ch.y : ch.x;
        bool entry = false, mid = false;
        if (c < nint) {
            const uint2 rc = rch[side];
            if (spans(rc, kSuperShift))
                entry = !cross || spans(rc, kHyperShift);
            else
                mid = spans(rc, kBlockShift);
        }
        if (entry) {
            ++pend;
        } else if (mid) {  // a phase-2 node: another XCD's L2 may hold the line
            ls.in[t][2 * side] = load_device_f4(tb.pcm + kPairStride * (size_t)p + side);
            ls.in[t][2 * side + 1] = load_device_f4(tb.paabb + 2 * (size_t)p + side);
        } else {
            ls.in[t][2 * side] = tb.pcm[kPairStride * (size_t)p + side];
            ls.in[t][2 * side + 1] = tb.paabb[2 * (size_t)p + side];
        }
    }
    ls.node[t] = p;
    ls.pw[t] = pw;
    ls.par[t] = q;
    ls.pend[t] = pend;
    ls.src[t] = 0u;
    ls.cnt[t] = 0u;
}

// Phases 2 and 3, S + 1 workgroups.  Workgroup sb < S resolves the phase-2
// list of super-block sb in LDS (or, longer than lds_limit, climbs from the
// phase-2 nodes' other children), then, behind a device-scope fence, sets
// its done flag.  Workgroup S resolves the top list at the same time: a
// climb that needs a record from a phase-2 workgroup waits for that
// super-block's flag and reads the record at device scope, so the top
// levels resolve while the lower super-blocks still run (the deep chains of
// a dense region are not on one path with every super-block).  A longer top
// list waits for every flag and is resolved in LDS part by part (one part
// per hyper-block, then the nodes spanning hyper-blocks); one with a part
// longer than kLdsList, or any with the climb test hook (lds_limit 0),
// climbs from the top nodes' other children.  Workgroup S, the last one
// dispatched, is the only one that waits, so every flag it waits for is
// set.  It then clears the counters and flags for the next build.
__global__ __launch_bounds__(kFillBlock) void fill_mid_kernel(int64_t n, TreeBufs tb,
                                                              uint32_t num_super,
                                                              uint32_t lds_limit) {
    __shared__ ListLds ls;
    constexpr uint32_t kMaxParts = 1024;  // hyper-blocks of trees below 2^30 leaves, + 1
    __shared__ uint32_t s_part_len[kMaxParts + 1];
    const uint32_t t = threadIdx.x;
    const uint32_t sb = blockIdx.x;
    const uint32_t nint = (uint32_t)(n - 1);
    const uint32_t S = num_super;
    if (sb < S) {
        const uint32_t *list2 = tb.lists + (size_t)sb * kSuper;
        const uint32_t m2 = tb.counts[list_count(sb)];
        if (m2 <= lds_limit) {
            if (t < m2) setup_entry(tb, ls, list2, t, nint, kBlockShift);
            __syncthreads();
            resolve_list<false>(tb, ls, m2);
        } else {
            // climb from every child of a phase-2 node that is not one itself
            const uint32_t lo = sb * (uint32_t)kSuper;
            const uint32_t hi = lo + kSuper - 1 < nint ? lo + kSuper - 1 : nint;
            for (uint32_t k = t; k < m2; k += kFillBlock) {
                const uint2 ch = tb.child[list2[k]];
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const uint32_t c = side ? ch.y : ch.x;
                    if (!(c < nint && spans(tb.range[c], kBlockShift))) climb_wg(tb, nint, c, lo, hi);
                }
            }
        }
        // every thread's stores have completed (reached L2) before the
        // fence writes the L2 back
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (t == 0) {
            __threadfence();
            __hip_atomic_store(tb.counts + done_flag(sb), 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    const uint32_t *top = tb.lists + (size_t)S * kSuper;
    const uint32_t m3 = tb.counts[list_count(S)];
    auto wait_all = [&]() {
        for (uint32_t k = t; k < S; k += kFillBlock)
            for (uint32_t spins = 0; __hip_atomic_load(tb.counts + done_flag(k), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) == 0u;
                 ++spins) {
                if (spins == kSpinLimit) {
                    atomicOr(tb.fault, kFaultWait);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        __syncthreads();
    };
    if (m3 <= lds_limit) {
        if (t < m3) setup_entry(tb, ls, top, t, nint, kSuperShift);
        __syncthreads();
        resolve_list<true>(tb, ls, m3);
        wait_all();
    } else {
        wait_all();
        __threadfence();  // acquire: the phase-2 records, past this XCD's L2
        // the parts' lengths: part h < H of the nodes in hyper-block h, part H
        const uint32_t H = (uint32_t)((n + (int64_t(1) << kHyperShift) - 1) >> kHyperShift);
        for (uint32_t h = t; h <= H; h += kFillBlock) s_part_len[h] = 0u;
        __syncthreads();
        auto part_of = [&](uint32_t node) {
            const uint2 r = tb.range[node];
            return spans(r, kHyperShift) ? H : (r.x >> kHyperShift);
        };
        for (uint32_t k = t; k < m3; k += kFillBlock) atomicAdd(&s_part_len[part_of(top[k])], 1u);
        __syncthreads();
        bool fits = lds_limit > 0u;
        for (uint32_t h = 0; h <= H; ++h) fits = fits && s_part_len[h] <= kLdsList;
        if (fits) {
            // one part at a time: its nodes listed (in the phase-2 lists'
            // space, read by now), each node's entry in tb.slot
            uint32_t *part = tb.lists;
            for (uint32_t h = 0; h <= H; ++h) {
                if (s_part_len[h] == 0u) continue;
                if (t == 0) ls.nstart = 0u;
                __syncthreads();
                for (uint32_t k = t; k < m3; k += kFillBlock) {
                    const uint32_t node = top[k];
                    if (part_of(node) != h) continue;
                    const uint32_t e = atomicAdd(&ls.nstart, 1u);
                    part[e] = node;
                    tb.slot[node] = e;
                }
                __syncthreads();
                const uint32_t mp = s_part_len[h];
                if (t < mp) setup_part_entry(tb, ls, part, t, nint, h == H);
                __syncthreads();
                resolve_list<false>(tb, ls, mp);
                __syncthreads();
            }
        } else {
        // climb from every child of a top node that is not a top node
        for (uint32_t k = t; k < m3; k += kFillBlock) {
            const uint2 ch = tb.child[top[k]];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t c = side ? ch.y : ch.x;
                if (!(c < nint && spans(tb.range[c], kSuperShift))) climb_wg(tb, nint, c, 0u, nint);
            }
        }
        }
    }
    __syncthreads();
    // the next build counts from zero; the next fill waits on fresh flags
    for (uint32_t k = t; k <= S; k += kFillBlock) {
        tb.counts[list_count(k)] = 0u;
        if (k < S) tb.counts[done_flag(k)] = 0u;
    }
}

}  // namespace

void launch_build_u32(const uint32_t *keys, int64_t n, const TreeBufs &t, const LeafPlace &lpl,
                      uint32_t *clear,
                      int64_t clear_n, uint32_t *slots, hipStream_t s) {
    if (n < 2) return;
    build_u32_kernel<<<(unsigned)((n - 1 + kBuildNodes - 1) / kBuildNodes), kBuildThreads, 0,
                       s>>>(keys, n, t, lpl, clear, clear_n, slots);
}

void launch_build_u64(const uint64_t *keys, int64_t n, const int64_t *sim_off, const TreeBufs &t,
                      const LeafPlace &lpl,
                      uint32_t *sim_root, uint32_t *clear, int64_t clear_n, uint32_t *slots,
                      hipStream_t s) {
    if (n < 2) return;
    build_u64_kernel<<<(unsigned)((n - 1 + kBuildNodes - 1) / kBuildNodes), kBuildThreads, 0,
                       s>>>(keys, n, sim_off, t, lpl, sim_root, clear, clear_n, slots);
}

void launch_fill(const uint32_t *idx, const float *mass, const float2 *pos, int64_t n,
                 const TreeBufs &t, bool climb, hipStream_t s) {
    if (n < 1) return;
    const unsigned blocks = (unsigned)((n + kFillBlock - 1) / kFillBlock);
    fill_local_kernel<<<blocks, kFillBlock, 0, s>>>(idx, mass, pos, n, t);
    if (n > kFillBlock) {
        const uint32_t num_super = num_supers(n);
        fill_mid_kernel<<<num_super + 1, kFillBlock, 0, s>>>(n, t, num_super,
                                                             climb ? 0u : kLdsList);
    }
}

int64_t fill_count_words(int64_t cap) { return list_count(num_supers(cap)) + 1; }
int64_t fill_list_words(int64_t cap) {
    return (int64_t)num_supers(cap) * (kSuper + kTopPer);
}

}  // namespace nbs
