// k_walk.hip -- Barnes-Hut tree walk with fused kick and drift.
//
// Reference: bh_vel_step_main (src/shaders/compute/barnes_hut.wgsl:48-111)
// walks the tree once per body with a private 256-entry stack, accepting a
// node when it is a leaf or length * rsqrt(d^2 + eps^2) < theta, otherwise
// pushing left then right (so right is visited first), and skipping the
// body's own leaf; then v += a*dt.  bh_pos_step_main (barnes_hut.wgsl:113-121)
// is a second dispatch doing x += v*dt.
//
// MI355X design: the 64 lanes of a wavefront hold 64 consecutive bodies in
// Morton order and share ONE depth-first stack in LDS whose entries are
// (node, lane mask).  A node is visited once per wavefront: its 16-byte
// {com, mass, length} record and 8-byte child pair are wave-uniform scalar
// loads, and each lane named in the mask applies the reference's own test.
// Lanes that open the node form the mask of its children.  Because every
// lane's accepted nodes, and the order it meets them, are exactly those of
// its private reference walk, each lane's sum is the reference's sum term
// for term.  The walk reads the tree only (the body's own position comes
// from its leaf record), so the drift is fused into the epilogue without
// changing any force: x_new = x + (v + a*dt)*dt, written by body index.
//
// Stack depth: keys are unique once tie-broken by position, the common
// prefix length grows by at least one per level, so depth <= 64 (96 for the
// ensemble's 38-bit keys) and the stack never holds more than depth + 1
// entries.  A full stack drops both children like the reference
// (barnes_hut.wgsl:100-105) and raises an overflow counter.
#include "walk_common.h"

namespace nbs {

namespace {

// Softening^2 of the reference's per-pair mode: eps = multiplier *
// sqrt(G * max(m1, m2) * dt) (barnes_hut.wgsl:84-86).
__device__ __forceinline__ float soft_eps2(float m1, float m2, const StepParams &p) {
    const float m_eff = fmaxf(m1, m2);
    const float eps = p.eps_mult * sqrtf(p.grav * m_eff * p.dt);
    return eps * eps;
}

// Adds the pull of node record c on a body at offset (rx, ry) from it,
// inv = 1/sqrt(|r|^2 + eps^2).  Exact mode: the reference's expression
// a + (G m) r inv^3 with every operation rounded as written
// (barnes_hut.wgsl:91-94).  Fast mode: a + m r inv^3 with fused
// multiply-adds; the caller applies G once to the sum (accel_out).
template <bool kExact>
__device__ __forceinline__ void add_pull(const float4 &c, float rx, float ry, float inv, float G,
                                         float &ax, float &ay) {
    if constexpr (kExact) {
        const float inv3 = inv * inv * inv;
        const float gm = G * c.z;
        ax = ax + gm * rx * inv3;
        ay = ay + gm * ry * inv3;
    } else {
        const float f = c.z * (inv * inv * inv);
        ax = __builtin_fmaf(f, rx, ax);
        ay = __builtin_fmaf(f, ry, ay);
    }
}

// The body's acceleration from the walk's sum (see add_pull).
template <bool kExact>
__device__ __forceinline__ float accel_out(float sum, float G) {
    if constexpr (kExact)
        return sum;
    else
        return G * sum;
}

// Bounding box of a lane's drifted positions as order-preserving images
// (float_order; 0 = none, numbers only, like fminf/fmaxf).
struct BoxAcc {
    uint32_t nx = 0, ny = 0, xx = 0, xy = 0;
    __device__ __forceinline__ void add(float2 x) {
        if (x.x == x.x) {
            nx = max(nx, ~float_order(x.x));
            xx = max(xx, float_order(x.x));
        }
        if (x.y == x.y) {
            ny = max(ny, ~float_order(x.y));
            xy = max(xy, float_order(x.y));
        }
    }
};

// Max over the wavefront (all lanes active): DPP inside each 16-lane row
// (lane ^ 1, lane ^ 2, the other quad, the other half-row), then the four
// row maxima; no LDS-crossbar permutes.
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = max(v, dpp_u32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = max(v, dpp_u32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = max(v, dpp_u32<0x141>(v));  // row_half_mirror
    v = max(v, dpp_u32<0x140>(v));  // row_mirror
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return max(max(r0, r1), max(r2, r3));
}

// Fold the wave's drifted positions into the bounding-box accumulators the
// next step's encode reads (one slot per wave index mod kBBoxSlots; an
// atomic max per non-empty component).  Called by whole wavefronts.
__device__ __forceinline__ void fold_bbox(uint32_t *slots, BoxAcc b, uint32_t wave) {
    if (!slots) return;
    b.nx = wave_max_u32(b.nx);
    b.ny = wave_max_u32(b.ny);
    b.xx = wave_max_u32(b.xx);
    b.xy = wave_max_u32(b.xy);
    if ((threadIdx.x & 63) == 0) {
        uint32_t *s = slots + 4 * (wave & (kBBoxSlots - 1));
        if (b.nx) atomicMax(s + 0, b.nx);
        if (b.ny) atomicMax(s + 1, b.ny);
        if (b.xx) atomicMax(s + 2, b.xx);
        if (b.xy) atomicMax(s + 3, b.xy);
    }
}

// Kick and drift of the body at sorted position t with acceleration
// (ax, ay) (barnes_hut.wgsl:110, 120), written by body index; returns the
// drifted position.
template <bool kExact>
__device__ __forceinline__ float2 kick_drift_v(const WalkArgs &a, uint32_t bi, float2 v, float px,
                                               float py, float ax, float ay) {
    const float dt = a.p.dt;
    float2 x;
    if constexpr (kExact) {
        v.x = v.x + ax * dt;  // barnes_hut.wgsl:110
        v.y = v.y + ay * dt;
        x.x = px + v.x * dt;  // barnes_hut.wgsl:120
        x.y = py + v.y * dt;
    } else {
        v.x = __builtin_fmaf(ax, dt, v.x);
        v.y = __builtin_fmaf(ay, dt, v.y);
        x.x = __builtin_fmaf(v.x, dt, px);
        x.y = __builtin_fmaf(v.y, dt, py);
    }
    a.vel[bi] = v;
    a.pos[bi] = x;
    if (a.accel) a.accel[bi] = make_float2(ax, ay);
    return x;
}
template <bool kExact>
__device__ __forceinline__ float2 kick_drift(const WalkArgs &a, uint32_t t, float px, float py,
                                             float ax, float ay) {
    const uint32_t bi = a.idx[t];
    return kick_drift_v<kExact>(a, bi, a.vel[bi], px, py, ax, ay);
}

// One interaction of body (px, py, m1) with node record c.  Returns true
// when the body must open the node; otherwise adds the node's pull to
// (ax, ay) (barnes_hut.wgsl:78-97).
template <bool kExact, bool kRefSoft>
__device__ __forceinline__ bool interact(const float4 c, bool leaf, float px, float py, float m1,
                                         const StepParams &p, float eps2c, float &ax, float &ay) {
    const float rx = c.x - px;
    const float ry = c.y - py;
    const float eps2 = kRefSoft ? soft_eps2(m1, c.z, p) : eps2c;
    float inv;
    const bool pass = probe_node<kExact>(c, rx, ry, eps2, p.theta, p.open_scale, p.open_bias, inv);
    if (leaf || pass) {
        add_pull<kExact>(c, rx, ry, inv, p.grav, ax, ay);
        return false;
    }
    return true;
}

// Record of node `node` for the walk start: the root record, or the child
// slot in its parent's pair (ensemble sub-tree roots).
__device__ __forceinline__ float4 node_record(const WalkArgs &a, uint32_t node) {
    if (node == 0u) return a.root[0];
    const uint32_t pw = a.parent[node];
    return a.pcm[pcm_slot(pw)];
}

// XCD-aware workgroup order.  The hardware deals workgroups round-robin over
// the 8 XCDs (workgroup b runs on XCD b % 8), and each XCD has its own L2.
// With run > 0, XCD x walks runs of `run` consecutive logical workgroups
// (Morton-consecutive bodies, so its L2 holds their region's sub-tree), the
// runs dealt to the XCDs in turn; run = 0 keeps the hardware order.  The
// grid is a multiple of 8 * run; logical workgroups past the bodies exit.
__device__ __forceinline__ uint32_t logical_block(uint32_t b, uint32_t run) {
    if (!run) return b;
    const uint32_t x = b & 7u, i = b >> 3;
    const uint32_t r = i / run;
    return (r * 8u + x) * run + (i - r * run);
}

// Grid and run length for `blocks` logical workgroups: xcd_order 0 =
// hardware order, 1 = one run per XCD, k > 1 = runs of k workgroups.
void xcd_grid(uint32_t blocks, int xcd_order, uint32_t &grid, uint32_t &run) {
    grid = blocks;
    run = 0;
    if (xcd_order <= 0 || blocks < 16) return;
    run = xcd_order == 1 ? (blocks + 7u) / 8u : (uint32_t)xcd_order;
    const uint32_t span = 8u * run;
    grid = (blocks + span - 1u) / span * span;
}

template <bool kExact, bool kRefSoft, bool kEnsemble, bool kStats>
__global__ __launch_bounds__(1024) void walk_kernel(const WalkArgs a, int stack_cap,
                                                    uint32_t xcd_run) {
    extern __shared__ float4 s_dyn[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int waves = blockDim.x >> 6;
    // LDS stack of one wavefront: (node record, lane mask, node id) entries.
    float4 *s_data = s_dyn + (size_t)wid * stack_cap;
    uint64_t *s_mask = reinterpret_cast<uint64_t *>(s_dyn + (size_t)waves * stack_cap) +
                       (size_t)wid * stack_cap;
    uint32_t *s_node = reinterpret_cast<uint32_t *>(reinterpret_cast<uint64_t *>(
                           s_dyn + (size_t)waves * stack_cap) + (size_t)waves * stack_cap) +
                       (size_t)wid * stack_cap;

    const uint32_t block = logical_block(blockIdx.x, xcd_run);

    const uint32_t n = a.n;
    const uint32_t t = block * blockDim.x + threadIdx.x;  // sorted position
    const bool valid = t < n;
    const uint32_t leaf0 = n - 1u;
    const uint32_t my_leaf = leaf0 + t;
    const_as<float> *pcmf_s = (const_as<float> *)a.pcm;
    const_as<uint32_t> *pcmu_s = (const_as<uint32_t> *)a.pcm;

    float px = 0.0f, py = 0.0f, m1 = 0.0f;
    if (valid) {
        const float4 me = a.leaf[t];  // written by the fill: {pos, mass, 0}
        px = me.x;
        py = me.y;
        m1 = me.z;
    }
    const uint64_t live = __ballot(valid);
    if (live == 0) return;

    const StepParams p = a.p;
    const float eps2c = p.eps * p.eps;
    float ax = 0.0f, ay = 0.0f;
    uint32_t n_acc = 0, n_open = 0, n_drop = 0;
    int sp = 0;
    uint32_t node = 0u;
    uint64_t mask = live;
    float4 c;
    if constexpr (kEnsemble) {
        // Each lane starts at its sim's sub-tree root; group lanes by root.
        uint32_t my_root = 0u;
        if (valid) my_root = a.sim_root[(uint32_t)(a.keys64[t] >> kSimShift)];
        uint64_t rem = live;
        bool first = true;
        while (rem) {
            const int l0 = __builtin_ctzll(rem);
            const uint32_t r = uniform_u32((uint32_t)__builtin_amdgcn_readlane((int)my_root, l0));
            const uint64_t m = __ballot(valid && my_root == r) & rem;
            rem &= ~m;
            const float4 rc = node_record(a, r);
            if (first) {
                node = r;
                mask = m;
                c = rc;
                first = false;
            } else if (sp < stack_cap) {
                if (lane == 0) {
                    s_node[sp] = r;
                    s_mask[sp] = m;
                    s_data[sp] = rc;
                }
                ++sp;
            } else if (lane == 0) {
                atomicOr(a.overflow, 1u);
            }
        }
    } else {
        c = a.root[0];
    }

    // A wavefront visits each node at most once per lane group, so a walk of a
    // valid tree ends within 4n iterations; the bound only guards against a
    // corrupt tree and reports like a stack overflow.
    const uint32_t max_iter = 4u * n + 64u;
    for (uint32_t iter = 0;; ++iter) {
        if (iter >= max_iter) {
            if (lane == 0) atomicOr(a.overflow, 1u);
            break;
        }
        const bool leaf = node >= leaf0;
        // Speculative scalar loads of the children (used if the node opens);
        // a leaf reads pair 0, which always exists when n >= 2.
        const size_t pn = leaf ? 0u : node;
        const_as<float> *rec = pcmf_s + 4 * kPairStride * pn;
        const float4 cl = make_float4(rec[0], rec[1], rec[2], rec[3]);
        const float4 cr = make_float4(rec[4], rec[5], rec[6], rec[7]);
        const uint32_t chl = pcmu_s[4 * kPairStride * pn + 8];
        const uint32_t chr = pcmu_s[4 * kPairStride * pn + 9];
        bool open = false;
        if (((mask >> lane) & 1u) && !(leaf && node == my_leaf)) {
            open = interact<kExact, kRefSoft>(c, leaf, px, py, m1, p, eps2c, ax, ay);
            if constexpr (kStats) {
                if (open)
                    ++n_open;
                else
                    ++n_acc;
            }
        }
        const uint64_t om = __ballot(open);
        if (om) {
            if (sp < stack_cap) {
                // Push left, continue with right: the reference pops right first.
                if (lane == 0) {
                    s_node[sp] = chl;
                    s_mask[sp] = om;
                    s_data[sp] = cl;
                }
                ++sp;
                node = chr;
                mask = om;
                c = cr;
                continue;
            }
            if (lane == 0) atomicOr(a.overflow, 1u);
            if constexpr (kStats) {
                if (open) {
                    --n_open;
                    ++n_drop;
                }
            }
        }
        if (sp == 0) break;
        --sp;
        node = uniform_u32(s_node[sp]);
        mask = uniform_u64(s_mask[sp]);
        c = s_data[sp];
    }

    BoxAcc box;
    if (valid)
        box.add(kick_drift<kExact>(a, t, px, py, accel_out<kExact>(ax, p.grav),
                                   accel_out<kExact>(ay, p.grav)));
    fold_bbox(a.bbox_slots, box, t >> 6);
    if constexpr (kStats) {
        // wave sums, one atomic per wave and counter
        uint32_t s0 = n_acc, s1 = n_open, s2 = n_drop;
        for (int o = 32; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
            s2 += __shfl_xor(s2, o);
        }
        if (lane == 0) {
            atomicAdd(&a.stats[0], (unsigned long long)s0);
            atomicAdd(&a.stats[1], (unsigned long long)s1);
            if (s2) atomicAdd(&a.stats[2], (unsigned long long)s2);
        }
    }
}

// Register-stack walk: each lane holds kB bodies, slot b of wavefront w
// being sorted position w * 64 * kB + 64 * b + lane, so the wavefront's
// 64 * kB Morton-consecutive bodies share one traversal (walk_common.h).
template <bool kExact, bool kRefSoft, bool kEnsemble, bool kStats, int kB>
__global__ __launch_bounds__(1024) void walk_rs_kernel(const WalkArgs a, int stack_cap,
                                                       uint32_t xcd_run) {
    const int lane = threadIdx.x & 63;
    const uint32_t block = logical_block(blockIdx.x, xcd_run);
    const uint32_t n = a.n;
    const uint32_t first = ((block * blockDim.x + threadIdx.x) >> 6) * (64u * kB);
    const uint32_t leaf0 = n - 1u;

    uint32_t t[kB], my_leaf[kB];
    bool valid[kB];
    float px[kB], py[kB], m1[kB], ax[kB], ay[kB];
    uint64_t mask[kB];
    uint64_t live = 0;
#pragma unroll
    for (int b = 0; b < kB; ++b) {
        t[b] = first + 64u * b + lane;  // sorted position
        valid[b] = t[b] < n;
        my_leaf[b] = leaf0 + t[b];
        px[b] = py[b] = m1[b] = ax[b] = ay[b] = 0.0f;
        if (valid[b]) {
            const float4 me = a.leaf[t[b]];  // written by the fill: {pos, mass, 0}
            px[b] = me.x;
            py[b] = me.y;
            m1[b] = me.z;
        }
        mask[b] = __ballot(valid[b]);
        live |= mask[b];
    }
    if (live == 0) return;

    const StepParams p = a.p;
    const float eps2c = p.eps * p.eps;
    RegStack<kB> st;
    int sp = 0;
    uint32_t node = 0u;
    float4 c;
    if constexpr (kEnsemble) {
        // Each body starts at its sim's sub-tree root; one entry per distinct
        // root, with the slot masks of the bodies under it.
        uint32_t my_root[kB];
        uint64_t rem[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            my_root[b] = valid[b] ? a.sim_root[(uint32_t)(a.keys64[t[b]] >> kSimShift)] : 0u;
            rem[b] = mask[b];
        }
        bool first = true;
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            while (rem[b]) {
                const int l0 = __builtin_ctzll(rem[b]);
                const uint32_t r =
                    uniform_u32((uint32_t)__builtin_amdgcn_readlane((int)my_root[b], l0));
                uint64_t m[kB];
#pragma unroll
                for (int b2 = 0; b2 < kB; ++b2) {
                    m[b2] = __ballot(valid[b2] && my_root[b2] == r) & rem[b2];
                    rem[b2] &= ~m[b2];
                }
                const float4 rc = node_record(a, r);
                if (first) {
                    node = r;
#pragma unroll
                    for (int b2 = 0; b2 < kB; ++b2) mask[b2] = m[b2];
                    c = rc;
                    first = false;
                } else if (sp < stack_cap) {
                    st.put(sp >> 6, sp & 63, r, m, rc);
                    ++sp;
                } else if (lane == 0) {
                    atomicOr(a.overflow, 1u);
                }
            }
        }
    } else {
        c = a.root[0];
    }

    // per-slot geometry of the node being visited, kept by probe for pull
    float rx[kB], ry[kB], inv[kB];
    uint64_t n_acc = 0, n_open = 0, n_drop = 0;
    walk_tree<kB>(
        a.pcm, leaf0, my_leaf, node, mask, c, st, sp, stack_cap, a.overflow,
        [&](const float4 &cc, int b) {
            rx[b] = cc.x - px[b];
            ry[b] = cc.y - py[b];
            const float eps2 = kRefSoft ? soft_eps2(m1[b], cc.z, p) : eps2c;
            return probe_node<kExact>(cc, rx[b], ry[b], eps2, p.theta, p.open_scale, p.open_bias,
                                      inv[b]);
        },
        [&](const float4 &cc, uint32_t, int b, uint64_t acc) {
            float nx = ax[b], ny = ay[b];
            add_pull<kExact>(cc, rx[b], ry[b], inv[b], p.grav, nx, ny);
            const bool in = __builtin_amdgcn_inverse_ballot_w64(acc);
            ax[b] = in ? nx : ax[b];
            ay[b] = in ? ny : ay[b];
        },
        [&](int, uint64_t acc, uint64_t open, uint64_t dropped) {
            if constexpr (kStats) {
                n_acc += __builtin_popcountll(acc);
                n_open += __builtin_popcountll(open);
                n_open -= __builtin_popcountll(dropped);
                n_drop += __builtin_popcountll(dropped);
            }
        });

    BoxAcc box;
#pragma unroll
    for (int b = 0; b < kB; ++b) {
        if (valid[b])
            box.add(kick_drift<kExact>(a, t[b], px[b], py[b], accel_out<kExact>(ax[b], p.grav),
                                       accel_out<kExact>(ay[b], p.grav)));
    }
    fold_bbox(a.bbox_slots, box, first >> 6);
    if constexpr (kStats) {
        // the counters are wave-uniform: one atomic per wave and counter
        if (lane == 0) {
            atomicAdd(&a.stats[0], (unsigned long long)n_acc);
            atomicAdd(&a.stats[1], (unsigned long long)n_open);
            if (n_drop) atomicAdd(&a.stats[2], (unsigned long long)n_drop);
        }
    }
}

// Pair walk (fast arithmetic, one body per lane).  The loop's unit is an
// opened node: its pair record holds both children's records and indices,
// so both children are tested in one iteration, from one scalar load, with
// two independent dependency chains.  The lanes that open a child form its
// mask; of two opened children the right one is walked next and the left
// one waits on the stack, as in the reference's walk (barnes_hut.wgsl:
// 100-105).  Every lane takes exactly the opening decisions of its private
// reference walk, with the fast mode's arithmetic (the same probe as the
// register-stack walk), so the accepted nodes, the visit counts and every
// term of the sum are the same; the order of the terms differs where a lane
// accepts a left child whose right sibling it opens (the pair walk adds it
// before the right subtree, the reference after), which the fast mode's
// tolerance covers.  A lane's order is fixed (a right-first pre-order of its
// own opened nodes), whatever other lanes share its wavefront.
//
// The loop is written for the scalar unit, which bounds it: the next
// record's load is issued before this node's pulls (their factors are
// computed first, so the load may overwrite the record's registers), the
// pulls are branch-free (a lane outside the accept mask keeps its sum), and
// every decision is an integer or a lane mask, never a branch on a boolean.
//
// kSelf: leave each body's own leaf out with a lane compare.  Without it the
// own leaf is accepted: its offset is exactly zero, so with a finite pull
// factor m * inv^3 (the host checks that eps^2 and the largest mass keep it
// finite) its term is an exact +0 and the sum is unchanged.
// Quadrupole expansion of a node's pull (central second moments Q = {xx, xy,
// yy} about its centre of mass, r = centre - body, inv = 1/sqrt(|r|^2 +
// eps^2)): the softened potential's Taylor series to second order gives
//   a = G [ (m inv^3 - k) r - 3 inv^5 Q r ],  k = inv^5 (1.5 tr Q - 7.5 inv^2 r.Q r),
// returned as the factor k1 = m inv^3 - k and the vector g = 3 inv^5 Q r
// (the sums take a += k1 r - g; G is applied once at the end).  A leaf has
// Q = 0: k1 = m inv^3 and g = 0 exactly.
struct QuadTerm {
    float k1, gx, gy;
};
__device__ __forceinline__ QuadTerm quad_term(float f0, float qxx, float qxy, float qyy, float rx,
                                              float ry, float inv) {
    const float inv2 = inv * inv;
    const float inv5 = (inv2 * inv2) * inv;
    const float qrx = __builtin_fmaf(qxx, rx, qxy * ry);
    const float qry = __builtin_fmaf(qxy, rx, qyy * ry);
    const float rqr = __builtin_fmaf(rx, qrx, ry * qry);
    const float k = inv5 * __builtin_fmaf(-7.5f * inv2, rqr, 1.5f * (qxx + qyy));
    const float g5 = 3.0f * inv5;
    return QuadTerm{f0 - k, g5 * qrx, g5 * qry};
}

// Second moments of node `node` (the root's, or from its parent's pair record).
__device__ __forceinline__ float3 node_quad(const WalkArgs &a, uint32_t node) {
    if (node == 0u) {
        const float4 q = a.root[2];
        return make_float3(q.x, q.y, q.z);
    }
    const uint32_t pw = a.parent[node];
    const float *rec = reinterpret_cast<const float *>(a.pcm + kPairStride * (size_t)(pw & kParentMask));
    const float *q = (pw >> 31) ? rec + 13 : rec + 10;
    return make_float3(q[0], q[1], q[2]);
}

// The pair walk of one chunk: the 64 bodies at sorted positions [64c, 64c + 64),
// one per lane of the calling wavefront.  kQuad: quadrupole expansion (the
// children's moments are the record's padding words).
template <bool kRefSoft, bool kEnsemble, bool kStats, bool kSelf, bool kOff32, bool kQuad>
__device__ __forceinline__ void pair_walk_chunk(const WalkArgs &a, uint32_t chunk) {
    const int lane = threadIdx.x & 63;
    const uint32_t n = a.n;
    // the chunk's bodies: wave_bodies (64, 32, 16 or 8) Morton-consecutive
    // bodies on the first lanes (fewer bodies per wavefront: more, shorter
    // traversals, for latency-bound small simulations; no sum changes)
    const uint32_t wb = a.wave_bodies;
    const uint32_t t = chunk * wb + (uint32_t)lane;  // sorted position
    const bool valid = (uint32_t)lane < wb && t < n;
    const uint32_t leaf0 = n - 1u;
    const uint32_t my_leaf = leaf0 + t;
    float px = 0.0f, py = 0.0f, m1 = 0.0f;
    uint32_t bi = 0u;  // the body, and its velocity for the epilogue, read
    float2 v0 = make_float2(0.0f, 0.0f);  // now: their latency hides behind the walk
    if (valid) {
        const float4 me = a.leaf[t];  // written by the build: {pos, mass, 0}
        bi = a.idx[t];
        v0 = a.vel[bi];
        px = me.x;
        py = me.y;
        m1 = me.z;
    }
    const uint64_t live = __ballot(valid);
    if (live == 0) return;

    const StepParams p = a.p;
    const float eps2c = p.eps * p.eps;
    float ax = 0.0f, ay = 0.0f;
    uint64_t n_acc = 0, n_open = 0, n_drop = 0;
    uint32_t my_acc = 0, my_open = 0;  // this lane's own counters (stats mode)
    auto own = [&](uint64_t m) { return (uint32_t)((m >> lane) & 1ull); };
    auto probe = [&](const float4 &c, float &rx, float &ry, float &inv) {
        rx = c.x - px;
        ry = c.y - py;
        const float eps2 = kRefSoft ? soft_eps2(m1, c.z, p) : eps2c;
        return probe_node<false>(c, rx, ry, eps2, p.theta, p.open_scale, p.open_bias, inv);
    };
    // a + f r for the lanes of `acc` (f = m inv^3: add_pull's fast form);
    // the other lanes add 0 * r, an exact zero for a finite offset (the sum
    // is never -0), so the factor is selected instead of both sums
    auto accumulate = [&](float f, float rx, float ry, uint64_t acc) {
        const float fs = __builtin_amdgcn_inverse_ballot_w64(acc) ? f : 0.0f;
        ax = __builtin_fmaf(fs, rx, ax);
        ay = __builtin_fmaf(fs, ry, ay);
    };
    // quadrupole form: a + k1 r - g for the lanes of `acc` (the others add
    // 0 r - 0, an exact zero)
    auto accumulate_q = [&](const QuadTerm &q, float rx, float ry, uint64_t acc) {
        const bool in = __builtin_amdgcn_inverse_ballot_w64(acc);
        const float ks = in ? q.k1 : 0.0f, gxs = in ? q.gx : 0.0f, gys = in ? q.gy : 0.0f;
        ax = __builtin_fmaf(ks, rx, ax) - gxs;
        ay = __builtin_fmaf(ks, ry, ay) - gys;
    };
    PairStack st;
    st.put(0, 0u, 0ull);  // the sentinel
    int sp = 1;
    // A start node (the root, or a sim's sub-tree root) is tested by its
    // lanes on its own record; the lanes that open it make a stack entry.
    auto start = [&](uint32_t r, uint64_t m, const float4 &c) {
        float rx, ry, inv;
        const uint64_t pass = __ballot(probe(c, rx, ry, inv));
        uint64_t acc = m & pass;
        // a start node that is a leaf is the root of a one-body (sub-)tree
        if (r >= leaf0) acc &= __ballot(my_leaf != r);
        const uint64_t open = m & ~pass;
        if (acc) {
            if constexpr (kQuad) {
                const float3 q = node_quad(a, r);
                accumulate_q(quad_term(c.z * (inv * inv * inv), q.x, q.y, q.z, rx, ry, inv), rx, ry,
                             acc);
            } else {
                accumulate(c.z * (inv * inv * inv), rx, ry, acc);
            }
        }
        if constexpr (kStats) {
            n_acc += __builtin_popcountll(acc);
            n_open += __builtin_popcountll(open);
            my_acc += own(acc);
            my_open += own(open);
        }
        if (open) {
            if (sp < kPairStackTop) {
                st.put(sp, r, open);
                ++sp;
            } else {
                *a.overflow = 1u;
                if constexpr (kStats) {
                    n_open -= __builtin_popcountll(open);
                    n_drop += __builtin_popcountll(open);
                    my_open -= own(open);
                }
            }
        }
    };
    if constexpr (kEnsemble) {
        const uint32_t my_root = valid ? a.sim_root[(uint32_t)(a.keys64[t] >> kSimShift)] : 0u;
        uint64_t rem = live;
        while (rem) {
            const int l0 = __builtin_ctzll(rem);
            const uint32_t r = uniform_u32((uint32_t)__builtin_amdgcn_readlane((int)my_root, l0));
            const uint64_t m = __ballot(valid && my_root == r) & rem;
            rem &= ~m;
            start(r, m, node_record(a, r));
        }
    } else {
        start(0u, live, a.root[0]);
    }

    uint32_t n_over = 0;  // dropped left children (stack full)
    if (sp > 1) {
        const_as<float> *pcm_s = (const_as<float> *)a.pcm;
        const_as<uint32_t> *pcmu_s = (const_as<uint32_t> *)a.pcm;
        uint32_t node;
        uint64_t mask;
        --sp;
        st.get(sp, node, mask);
        // the pair record of `node` (dwords: HIP vector types cannot be
        // loaded through the constant address space); 64-bit offsets.  The
        // first one passes through readfirstlane: a loop-entry value that is
        // a load of the same form as the loop's own would let the compiler
        // merge both loads into one at the loop head, where nothing can hide
        // its latency.
        size_t rb = 4 * kPairStride * (size_t)node;
        float4 cl = make_float4(uniform_f32(pcm_s[rb + 0]), uniform_f32(pcm_s[rb + 1]),
                                uniform_f32(pcm_s[rb + 2]), uniform_f32(pcm_s[rb + 3]));
        float4 cr = make_float4(uniform_f32(pcm_s[rb + 4]), uniform_f32(pcm_s[rb + 5]),
                                uniform_f32(pcm_s[rb + 6]), uniform_f32(pcm_s[rb + 7]));
        uint2 ch = make_uint2(uniform_u32(pcmu_s[rb + 8]), uniform_u32(pcmu_s[rb + 9]));
        // the first record's padding: the children's opening keys (monopole)
        // or moments (quadrupole)
        PairPad pad = {};
#pragma unroll
        for (int k = 0; k < (kQuad ? 6 : 2); ++k) pad.w[k] = uniform_u32(pcmu_s[rb + 10 + k]);
        // The pull of one child for the lanes of `acc` (non-empty), from its
        // mass, moments, offset and squared softened distance d2: the
        // reciprocal square root is taken only here, for the children some
        // lane accepts (about a quarter of the children no lane accepts).
        auto pull = [&](float m, uint32_t q0, uint32_t q1, uint32_t q2, float rx, float ry,
                        float d2, uint64_t acc) {
            const float inv = __builtin_amdgcn_rsqf(d2);
            const float f = m * (inv * inv * inv);
            if constexpr (kQuad)
                accumulate_q(quad_term(f, __uint_as_float(q0), __uint_as_float(q1),
                                       __uint_as_float(q2), rx, ry, inv),
                             rx, ry, acc);
            else
                accumulate(f, rx, ry, acc);
        };
        for (;;) {
            keep_live(pad);  // one 64-byte load per record
            const float eps2l = kRefSoft ? soft_eps2(m1, cl.z, p) : eps2c;
            const float eps2r = kRefSoft ? soft_eps2(m1, cr.z, p) : eps2c;
            const float rxl = cl.x - px, ryl = cl.y - py, rxr = cr.x - px, ryr = cr.y - py;
            const float d2l = soft_d2(rxl, ryl, eps2l), d2r = soft_d2(rxr, ryr, eps2r);
            // the opening keys: the record's own (monopole), or computed from
            // the lengths with the same expression (quadrupole: the padding
            // holds moments)
            float kl, kr;
            if constexpr (kQuad) {
                kl = open_key(cl.w, p.open_scale, p.open_bias);
                kr = open_key(cr.w, p.open_scale, p.open_bias);
            } else {
                kl = __uint_as_float(pad.w[0]);
                kr = __uint_as_float(pad.w[1]);
            }
            const uint64_t bl = __ballot(passes_key(kl, d2l));
            const uint64_t br = __ballot(passes_key(kr, d2r));
            // leaves always pass, so an opened child is an internal node.
            // The next entry: the right child first, the left one waits on
            // the stack (written to the free slot every time, kept by
            // advancing the top); with neither child opened, the top entry
            // (the sentinel's empty mask ends the walk).  A plain tree of up
            // to 2^26 bodies is at most 32 + 26 levels deep (30-bit keys tie-
            // broken on the sorted position), so its stack never reaches the
            // top; ensembles' 38-bit keys and larger trees keep the clamp.
            constexpr bool kClamp = kEnsemble || kStats || !kOff32;
            uint32_t next, sp_next, both;
            uint64_t accl, accr, openl, openr;
            pair_next<kClamp>(mask, bl, br, ch.x, ch.y, (uint32_t)sp, accl, accr, openl, openr,
                              next, sp_next, both);  // mask: the next entry's
            if constexpr (kSelf) {
                accl &= __ballot(my_leaf != ch.x);
                accr &= __ballot(my_leaf != ch.y);
            }
            if constexpr (kStats) {
                n_acc += __builtin_popcountll(accl) + __builtin_popcountll(accr);
                n_open += __builtin_popcountll(openl) + __builtin_popcountll(openr);
                my_acc += own(accl) + own(accr);
                my_open += own(openl) + own(openr);
            }
            st.put(sp, ch.x, openl);
            if constexpr (kClamp) {
                // both children opened with the top at kPairStackTop: the
                // left child is dropped
                const uint32_t drop = both & (uint32_t)(sp == kPairStackTop);
                n_over += drop;
                if constexpr (kStats) {
                    if (drop) {
                        n_open -= __builtin_popcountll(openl);
                        n_drop += __builtin_popcountll(openl);
                        my_open -= own(openl);
                    }
                }
            }
            sp = (int)sp_next;
            if (mask == 0) st.get(sp, next, mask);
            // what the pulls take from this record, copied out of the
            // record's registers so that the next record lands in them whole
            const float ml = scalar_copy(cl.z), mr = scalar_copy(cr.z);
            PairPad q;
            if constexpr (kQuad) {
#pragma unroll
                for (int k = 0; k < 6; ++k) q.w[k] = scalar_copy_u32(pad.w[k]);
            }
            // the next record, one 64-byte load, issued before this node's pulls
            {
                const_as<RecordWords> *rf;
                if constexpr (kOff32)  // byte offset below 2^32 (n <= 2^26): an SGPR offset
                    rf = (const_as<RecordWords> *)((const_as<char> *)pcm_s + (uint32_t)(next * 64u));
                else
                    rf = (const_as<RecordWords> *)(pcm_s + 4 * kPairStride * (size_t)next);
                const RecordWords w = *rf;
                cl = make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]),
                                 __uint_as_float(w[2]), __uint_as_float(w[3]));
                cr = make_float4(__uint_as_float(w[4]), __uint_as_float(w[5]),
                                 __uint_as_float(w[6]), __uint_as_float(w[7]));
                ch = make_uint2(w[8], w[9]);
#pragma unroll
                for (int k = 0; k < 6; ++k) pad.w[k] = w[10 + k];
            }
            if (accr) pull(mr, q.w[3], q.w[4], q.w[5], rxr, ryr, d2r, accr);
            if (accl) pull(ml, q.w[0], q.w[1], q.w[2], rxl, ryl, d2l, accl);
            if (mask == 0) break;
        }
    }
    if (n_over) *a.overflow = 1u;

    BoxAcc box;
    if (valid)
        box.add(kick_drift_v<false>(a, bi, v0, px, py, accel_out<false>(ax, p.grav),
                                    accel_out<false>(ay, p.grav)));
    fold_bbox(a.bbox_slots, box, t >> 6);
    if constexpr (kStats) {
        // the counters are wave-uniform: one atomic per wave and counter
        if (lane == 0) {
            atomicAdd(&a.stats[0], (unsigned long long)n_acc);
            atomicAdd(&a.stats[1], (unsigned long long)n_open);
            if (n_drop) atomicAdd(&a.stats[2], (unsigned long long)n_drop);
        }
        if (valid && a.body_counts) a.body_counts[a.idx[t]] = make_uint2(my_acc, my_open);
    }
}

template <bool kRefSoft, bool kEnsemble, bool kStats, bool kSelf, bool kOff32, bool kQuad>
__global__ __launch_bounds__(1024) void walk_pair_kernel(const WalkArgs a, uint32_t xcd_run) {
    const uint32_t block = logical_block(blockIdx.x, xcd_run);
    pair_walk_chunk<kRefSoft, kEnsemble, kStats, kSelf, kOff32, kQuad>(
        a, (block * blockDim.x + threadIdx.x) >> 6);
}

// Persistent work-queue walk: `resident` wavefronts (as many as the GPU
// holds at once) take chunks from a queue, each its own index first, then
// resident + the next value of the queue counter (zeroed by the build with
// the sort's scratch), until the chunks run out; the queue lists the chunks
// in a.chunk_order (null: sorted order).  Ensembles list their largest
// simulations' chunks first, so the longest work items start early and the
// short ones fill the end of the launch.
template <bool kRefSoft, bool kEnsemble, bool kStats, bool kSelf, bool kOff32, bool kQuad>
__global__ __launch_bounds__(256) void walk_pair_queue_kernel(const WalkArgs a, uint32_t resident) {
    const uint32_t chunks = (a.n + a.wave_bodies - 1u) / a.wave_bodies;
    uint32_t item = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    while (item < chunks) {
        const uint32_t c = a.chunk_order ? a.chunk_order[item] : item;
        pair_walk_chunk<kRefSoft, kEnsemble, kStats, kSelf, kOff32, kQuad>(a, uniform_u32(c));
        uint32_t next = 0;
        if ((threadIdx.x & 63) == 0) next = atomicAdd(a.queue, 1u);
        item = resident + uniform_u32(next);  // lane 0's value
    }
}

// Register-stack capacity for a requested stack size: one or two rows of 64.
int rs_stack_cap(int stack) { return stack < 64 ? 64 : (stack > 128 ? 128 : stack); }

template <bool kRefSoft, bool kEnsemble, bool kStats, bool kSelf, bool kQuad = false>
void launch_pair(const WalkArgs &a, int block, int xcd, bool queue, hipStream_t s) {
    const uint32_t chunks = (a.n + a.wave_bodies - 1u) / a.wave_bodies;
    if (queue && a.queue && a.resident) {
        const uint32_t waves = chunks < a.resident ? chunks : a.resident;
        const uint32_t grid = (waves + 3u) / 4u;  // 4 waves per workgroup
        const uint32_t resident = grid * 4u;
        if (a.n <= (1u << 26))
            walk_pair_queue_kernel<kRefSoft, kEnsemble, kStats, kSelf, true, kQuad>
                <<<grid, 256, 0, s>>>(a, resident);
        else
            walk_pair_queue_kernel<kRefSoft, kEnsemble, kStats, kSelf, false, kQuad>
                <<<grid, 256, 0, s>>>(a, resident);
        return;
    }
    uint32_t grid, run;
    const uint32_t waves_per_block = (uint32_t)block / 64u;
    xcd_grid((chunks + waves_per_block - 1u) / waves_per_block, xcd, grid, run);
    if (a.n <= (1u << 26))
        walk_pair_kernel<kRefSoft, kEnsemble, kStats, kSelf, true, kQuad>
            <<<grid, block, 0, s>>>(a, run);
    else
        walk_pair_kernel<kRefSoft, kEnsemble, kStats, kSelf, false, kQuad>
            <<<grid, block, 0, s>>>(a, run);
}

template <bool kExact, bool kRefSoft, bool kEnsemble, bool kStats, int kB>
void launch_rs(const WalkArgs &a, int block, int stack, int xcd, hipStream_t s) {
    const uint32_t per_block = (uint32_t)block * kB;
    uint32_t grid, run;
    xcd_grid((a.n + per_block - 1u) / per_block, xcd, grid, run);
    walk_rs_kernel<kExact, kRefSoft, kEnsemble, kStats, kB>
        <<<grid, block, 0, s>>>(a, rs_stack_cap(stack), run);
}

template <bool kExact, bool kRefSoft, bool kEnsemble, bool kStats>
void launch_t(const WalkArgs &a, int block, int stack, int xcd, int variant, int lane_bodies,
              hipStream_t s) {
    if constexpr (!kExact && !kRefSoft) {
        // the quadrupole expansion (constant softening, fast arithmetic) runs
        // a pair walk whatever the variant
        if (a.p.expansion == 1) {
            const bool queue = variant == 4;
            if constexpr (kStats) {
                launch_pair<false, kEnsemble, true, true, true>(a, block, xcd, queue, s);
            } else {
                if (a.self_zero)
                    launch_pair<false, kEnsemble, false, false, true>(a, block, xcd, queue, s);
                else
                    launch_pair<false, kEnsemble, false, true, true>(a, block, xcd, queue, s);
            }
            return;
        }
    }
    if constexpr (!kExact) {
        if (variant == 2 || variant == 4) {
            // the own-leaf compare is needed for the counters and whenever
            // the own leaf's term might not be an exact zero
            const bool queue = variant == 4;
            if constexpr (kStats || kRefSoft) {
                launch_pair<kRefSoft, kEnsemble, kStats, true>(a, block, xcd, queue, s);
            } else {
                if (a.self_zero)
                    launch_pair<kRefSoft, kEnsemble, kStats, false>(a, block, xcd, queue, s);
                else
                    launch_pair<kRefSoft, kEnsemble, kStats, true>(a, block, xcd, queue, s);
            }
            return;
        }
    }
    // the pair walks sum in their own order: the exact mode, which keeps the
    // reference's order, runs the register-stack walk instead
    if (variant >= 2) variant = 1;
    if (variant == 1) {
        if (lane_bodies == 4)
            launch_rs<kExact, kRefSoft, kEnsemble, kStats, 4>(a, block, stack, xcd, s);
        else if (lane_bodies == 2)
            launch_rs<kExact, kRefSoft, kEnsemble, kStats, 2>(a, block, stack, xcd, s);
        else
            launch_rs<kExact, kRefSoft, kEnsemble, kStats, 1>(a, block, stack, xcd, s);
        return;
    }
    uint32_t grid, run;
    xcd_grid((a.n + (unsigned)block - 1u) / (unsigned)block, xcd, grid, run);
    const size_t lds = (size_t)(block / kWave) * (size_t)stack *
                       (sizeof(float4) + sizeof(uint64_t) + sizeof(uint32_t));
    walk_kernel<kExact, kRefSoft, kEnsemble, kStats><<<grid, block, lds, s>>>(a, stack, run);
}

template <bool kExact, bool kRefSoft>
void launch_e(const WalkArgs &a, bool ens, bool stats, int block, int stack, int xcd,
              int variant, int lb, hipStream_t s) {
    if (ens) {
        if (stats)
            launch_t<kExact, kRefSoft, true, true>(a, block, stack, xcd, variant, lb, s);
        else
            launch_t<kExact, kRefSoft, true, false>(a, block, stack, xcd, variant, lb, s);
    } else {
        if (stats)
            launch_t<kExact, kRefSoft, false, true>(a, block, stack, xcd, variant, lb, s);
        else
            launch_t<kExact, kRefSoft, false, false>(a, block, stack, xcd, variant, lb, s);
    }
}

}  // namespace

void launch_walk(const WalkArgs &a, bool ensemble, bool stats, int block, int stack,
                 int xcd_order, int variant, int lane_bodies, hipStream_t s) {
    if (a.n == 0) return;
    const bool exact = a.p.math_mode != 0;
    const bool refsoft = a.p.soft_mode == 0;
    if (exact) {
        if (refsoft)
            launch_e<true, true>(a, ensemble, stats, block, stack, xcd_order, variant, lane_bodies, s);
        else
            launch_e<true, false>(a, ensemble, stats, block, stack, xcd_order, variant, lane_bodies, s);
    } else {
        if (refsoft)
            launch_e<false, true>(a, ensemble, stats, block, stack, xcd_order, variant, lane_bodies, s);
        else
            launch_e<false, false>(a, ensemble, stats, block, stack, xcd_order, variant, lane_bodies, s);
    }
}

}  // namespace nbs
