// walk_common.h -- the wavefront-coherent tree traversal shared by the
// forward walk (k_walk.hip) and the adjoint walk (k_adjoint.hip).
//
// The lanes of a wavefront hold 64 * kB consecutive Morton-ordered bodies
// (kB per lane) and walk one depth-first traversal together.  A traversal
// entry is (node, node record, one lane mask per body slot); every body in
// a mask applies its own opening test, and the bodies that open a node form
// the masks of its children.
// Right children are visited before left ones, as the reference's stack pops
// them (barnes_hut.wgsl:100-105), so every lane meets exactly the nodes of its
// private reference walk, in the same order.
#pragma once

#include "nbody_internal.h"

// v_writelane_b32 through its LLVM intrinsic (no clang builtin in this toolchain).
__device__ int nbs_writelane(int value, int lane_sel, int old) __asm("llvm.amdgcn.writelane.i32");

namespace nbs {

// Tree records are read-only during a walk and indexed by a wave-uniform
// node id; reading them through the constant address space makes them
// scalar loads (s_load) issued once per wavefront.
template <typename T>
using const_as = const __attribute__((address_space(4))) T;

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = uniform_u32((uint32_t)v), hi = uniform_u32((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Stack of the pair walk: entry k = (node, lane mask) in lane k of three
// VGPRs (see RegStack).  The node's record is not kept: an entry is a node
// that some lanes open, and what the walk needs next is its pair record,
// loaded when the entry is popped.  Lane 0 holds a sentinel (node 0, empty
// mask): popping it ends the walk, so a pop needs no emptiness test and the
// popped node always has a record to load.  Lane 63 stays free: the walk
// writes the slot above the top on every iteration and advances the top only
// when it pushes, which keeps the three registers in place (a conditional
// write makes the compiler copy them on every iteration).  The stack holds at
// most one pending left child per level of the current path, so its 62
// entries serve any tree of depth up to 63 (30-bit keys tie-broken on
// positions below 2^30 stay below that; deeper trees raise the overflow flag).
constexpr int kPairStackTop = 63;  // highest top (lanes 0..62 in use)

// 1 when the wave-uniform mask m is not empty, computed on the scalar unit
// (written as C++, the compiler turns the flag into a vector select and a
// readfirstlane).  An ALU compare and select only: no memory access.
__device__ __forceinline__ uint32_t nonzero_flag(uint64_t m) {
    uint32_t r;
    asm volatile("s_cmp_lg_u64 %1, 0\n\ts_cselect_b32 %0, 1, 0" : "=s"(r) : "s"(m) : "scc");
    return r;
}
// The pair walk's step on the scalar unit: from the visiting lanes `mask`
// and the lanes that accept the left and right child (bl, br), the lanes
// that accept (accl, accr) and open (openl, openr) each child, and the next
// entry, written over `mask`: the right child (ch_r, with its open mask) when
// some lane opens it, else the left one, else the stack's top entry (mask is
// then 0 and the caller pops entry sp_next).  The caller writes (ch_l, openl)
// to slot sp, which becomes an entry when both children are opened: sp_next =
// sp + 1 (push), sp (one child) or sp - 1 (pop).  With kClamp the top stays
// at kPairStackTop (a push there lands on the free slot and the left child is
// dropped: both = 1 reports it); without, the caller guarantees a tree less
// than kPairStackTop levels deep.  Straight-line scalar ALU code on the flags
// the mask operations set (written in C++, each select becomes a lane-mask
// materialisation of several instructions).
template <bool kClamp>
__device__ __forceinline__ void pair_next(uint64_t &mask, uint64_t bl, uint64_t br, uint32_t ch_l,
                                          uint32_t ch_r, uint32_t sp, uint64_t &accl,
                                          uint64_t &accr, uint64_t &openl, uint64_t &openr,
                                          uint32_t &next, uint32_t &sp_next, uint32_t &both) {
    // delta = R ? L : L - 1 with L, R the two flags (1: some lane opens)
    uint32_t hl, hlm1, delta;
#define NBS_PAIR_NEXT_HEAD                     \
    "s_and_b64 %[al], %[m], %[bl]\n\t"        \
    "s_and_b64 %[ar], %[m], %[br]\n\t"        \
    "s_andn2_b64 %[opl], %[m], %[bl]\n\t"     \
    "s_cselect_b32 %[hl], 1, 0\n\t"           \
    "s_add_i32 %[hlm1], %[hl], -1\n\t"        \
    "s_andn2_b64 %[opr], %[m], %[br]\n\t"     \
    "s_cselect_b64 %[m], %[opr], %[opl]\n\t"  \
    "s_cselect_b32 %[nx], %[cr], %[cl]\n\t"   \
    "s_cselect_b32 %[dl], %[hl], %[hlm1]\n\t"
#define NBS_PAIR_NEXT_OUT                                                                  \
    [m] "+s"(mask), [al] "=&s"(accl), [ar] "=&s"(accr), [opl] "=&s"(openl),              \
        [opr] "=&s"(openr), [nx] "=&s"(next), [hl] "=&s"(hl), [hlm1] "=&s"(hlm1),          \
        [dl] "=&s"(delta), [spn] "=&s"(sp_next)
#define NBS_PAIR_NEXT_IN \
    [bl] "s"(bl), [br] "s"(br), [cl] "s"(ch_l), [cr] "s"(ch_r), [sp] "s"(sp)
    if constexpr (kClamp) {
        asm volatile(NBS_PAIR_NEXT_HEAD
                     "s_cselect_b32 %[both], %[hl], 0\n\t"
                     "s_add_u32 %[spn], %[sp], %[dl]"
                     : NBS_PAIR_NEXT_OUT, [both] "=&s"(both)
                     : NBS_PAIR_NEXT_IN
                     : "scc");
        sp_next = min(sp_next, (uint32_t)kPairStackTop);
    } else {
        asm volatile(NBS_PAIR_NEXT_HEAD "s_add_u32 %[spn], %[sp], %[dl]"
                     : NBS_PAIR_NEXT_OUT
                     : NBS_PAIR_NEXT_IN
                     : "scc");
        both = 0u;
    }
#undef NBS_PAIR_NEXT_HEAD
#undef NBS_PAIR_NEXT_OUT
#undef NBS_PAIR_NEXT_IN
}
__device__ __forceinline__ float uniform_f32(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// A pair record is 64 bytes: two child records, the two child indices and six
// pad dwords.  Loading only the ten used dwords makes an s_load_dwordx8 and
// an s_load_dwordx2 of the same line, two requests to the scalar cache where
// one does, and the pair walks are bound by those requests (one load per
// record: walk -3.5 %).  The loop loads the pad dwords too, and keep_live()
// gives them a use at the loop head, so the record stays one s_load_dwordx16.
struct PairPad {
    uint32_t w[6];
};
// A whole pair record as one 16-dword scalar load.
typedef uint32_t RecordWords __attribute__((ext_vector_type(16)));
// A wave-uniform value moved to a register of its own (s_mov_b32), so that
// the register it came from can take a new load while the copy stays in use.
__device__ __forceinline__ uint32_t scalar_copy_u32(uint32_t v) {
    uint32_t r;
    asm volatile("s_mov_b32 %0, %1" : "=s"(r) : "s"(v));
    return r;
}
__device__ __forceinline__ float scalar_copy(float v) {
    return __uint_as_float(scalar_copy_u32(__float_as_uint(v)));
}
__device__ __forceinline__ void load_pad(PairPad &p, const_as<uint32_t> *ru) {
#pragma unroll
    for (int k = 0; k < 6; ++k) p.w[k] = ru[10 + k];
}
__device__ __forceinline__ void keep_live(const PairPad &p) {
    asm volatile("" ::"s"(p.w[0]), "s"(p.w[1]), "s"(p.w[2]), "s"(p.w[3]), "s"(p.w[4]),
                 "s"(p.w[5]));
}
struct PairStack {
    int node, mlo, mhi;
    __device__ __forceinline__ void put(int l, uint32_t nd, uint64_t m) {
        node = nbs_writelane((int)nd, l, node);
        mlo = nbs_writelane((int)(uint32_t)m, l, mlo);
        mhi = nbs_writelane((int)(uint32_t)(m >> 32), l, mhi);
    }
    __device__ __forceinline__ void get(int l, uint32_t &nd, uint64_t &m) const {
        nd = (uint32_t)__builtin_amdgcn_readlane(node, l);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(mlo, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(mhi, l);
        m = ((uint64_t)hi << 32) | lo;
    }
};

// Register-resident stack of one wavefront: entry k lives in lane k & 63 of
// row k >> 6, one VGPR per field, so a push or pop needs no LDS access, no
// exec-mask change and no readfirstlane.  A push writes lane k of each field
// with v_writelane_b32 (the backend puts the lane select in M0 as GFX9
// requires), a pop reads it back with v_readlane_b32.  No memory is accessed.
// An entry is (node, node record, one lane mask per body slot).
template <int kB>
struct RegStack {
    int node[2], cx[2], cy[2], cm[2], cw[2];
    int mlo[kB][2], mhi[kB][2];

    template <int r>
    __device__ __forceinline__ void put_row(int l, uint32_t nd, const uint64_t (&m)[kB],
                                            const float4 &c) {
        node[r] = nbs_writelane((int)nd, l, node[r]);
        cx[r] = nbs_writelane(__float_as_int(c.x), l, cx[r]);
        cy[r] = nbs_writelane(__float_as_int(c.y), l, cy[r]);
        cm[r] = nbs_writelane(__float_as_int(c.z), l, cm[r]);
        cw[r] = nbs_writelane(__float_as_int(c.w), l, cw[r]);
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            mlo[b][r] = nbs_writelane((int)(uint32_t)m[b], l, mlo[b][r]);
            mhi[b][r] = nbs_writelane((int)(uint32_t)(m[b] >> 32), l, mhi[b][r]);
        }
    }
    __device__ __forceinline__ void put(int row, int l, uint32_t nd, const uint64_t (&m)[kB],
                                        const float4 &c) {
        if (row == 0)
            put_row<0>(l, nd, m, c);
        else
            put_row<1>(l, nd, m, c);
    }
    template <int r>
    __device__ __forceinline__ void get_row(int l, uint32_t &nd, uint64_t (&m)[kB],
                                            float4 &c) const {
        nd = (uint32_t)__builtin_amdgcn_readlane(node[r], l);
        c.x = __int_as_float(__builtin_amdgcn_readlane(cx[r], l));
        c.y = __int_as_float(__builtin_amdgcn_readlane(cy[r], l));
        c.z = __int_as_float(__builtin_amdgcn_readlane(cm[r], l));
        c.w = __int_as_float(__builtin_amdgcn_readlane(cw[r], l));
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(mlo[b][r], l);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(mhi[b][r], l);
            m[b] = ((uint64_t)hi << 32) | lo;
        }
    }
    __device__ __forceinline__ void get(int row, int l, uint32_t &nd, uint64_t (&m)[kB],
                                        float4 &c) const {
        if (row == 0)
            get_row<0>(l, nd, m, c);
        else
            get_row<1>(l, nd, m, c);
    }
};

// Distance term of one lane: lengths, distances and 1/sqrt exactly as the
// forward walk computes them (barnes_hut.wgsl:81-94), so every walk built on
// it takes the same opening decisions.  Returns inv = 1/sqrt(|r|^2 + eps^2).
// The fast arithmetic's squared softened distance |r|^2 + eps^2.
__device__ __forceinline__ float soft_d2(float rx, float ry, float eps2) {
    return __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, eps2));
}
template <bool kExact>
__device__ __forceinline__ float inv_dist(float rx, float ry, float eps2) {
    if constexpr (kExact) {
        const float d2 = rx * rx + ry * ry;
        return 1.0f / sqrtf(d2 + eps2);
    } else {
        return __builtin_amdgcn_rsqf(soft_d2(rx, ry, eps2));
    }
}

// The reference's opening test, length * inv < theta (barnes_hut.wgsl:89),
// with its leaf rule (a leaf is accepted whatever the test says,
// barnes_hut.wgsl:94): a leaf's length is kLeafLength, the only negative
// one.  A product that is not a number (an internal node of length 0 at
// distance 0, or theta NaN) fails, so the node opens, as in the reference.
// The exact arithmetic's test.
__device__ __forceinline__ bool passes(float length, float inv, float theta) {
    return length * inv < theta || length < 0.0f;
}
// The fast arithmetic's test on a node's opening key (open_key) and the
// squared softened distance d2: every leaf passes (key -inf or NaN).
__device__ __forceinline__ bool passes_key(float key, float d2) { return !(key >= d2); }

// One body's opening test of node record c at offset (rx, ry): true when the
// body accepts the node; inv = 1/sqrt(|r|^2 + eps^2) for its pull.  Every
// walk of one arithmetic mode decides through this function (the pair walks
// read the same keys from their records), so all of them, forward and
// adjoint, take the same decisions.
template <bool kExact>
__device__ __forceinline__ bool probe_node(const float4 &c, float rx, float ry, float eps2,
                                           float theta, float open_scale, float open_bias,
                                           float &inv) {
    if constexpr (kExact) {
        inv = inv_dist<true>(rx, ry, eps2);
        return passes(c.w, inv, theta);
    } else {
        const float d2 = soft_d2(rx, ry, eps2);
        inv = __builtin_amdgcn_rsqf(d2);
        return passes_key(open_key(c.w, open_scale, open_bias), d2);
    }
}

// Depth-first traversal of one wavefront from (node, mask, c) with `sp`
// entries already on `st`.  Each lane holds kB bodies; slot b's body in this
// lane owns leaf my_leaf[b] and mask[b] names the slot's bodies that visit
// the node.  Per visited node and slot:
//   probe(c, b) -> bool: passes(c.w, inv, theta) for the lane's body; it
//                  also keeps what pull needs (called by the whole wave);
//   pull(c, node, b, acc): the bodies in lane mask `acc` (non-empty) accept;
//   tally(b, acc, open, dropped): lane masks for statistics.
// The bodies that pass accept the node (a leaf always passes; its owner is
// left out), the others open it and form the mask of its children.  Leaf
// children are taken at once (they have no subtree); of two internal
// children the right one is walked first and the left one waits on the
// stack, as the reference's stack pops them (barnes_hut.wgsl:100-105).  All
// mask logic is scalar.  The stack holds stack_cap entries, at least 64 (the
// tree depth, and so the stack depth, stays below 64 for 32-bit keys); past
// stack_cap the left child is dropped and *overflow raised (the reference
// drops both children of a full stack; either result is flagged as wrong).
// A wavefront visits a node at most once, so a tree from the build kernels
// ends within 2 * leaf0 + 1 iterations.  There is no iteration counter: its
// compare-and-exit costs 12-19% of the walk time at 800k bodies.
template <int kB, typename Probe, typename Pull, typename Tally>
__device__ __forceinline__ void walk_tree(const float4 *pcm, uint32_t leaf0,
                                          const uint32_t (&my_leaf)[kB], uint32_t node,
                                          uint64_t (&mask)[kB], float4 c, RegStack<kB> &st,
                                          int sp, int stack_cap, uint32_t *overflow,
                                          Probe &&probe, Pull &&pull, Tally &&tally) {
    const_as<float> *pcm_s = (const_as<float> *)pcm;
    const_as<uint32_t> *pcmu_s = (const_as<uint32_t> *)pcm;
    // The bodies of `m` that accept leaf `leaf_node` (all but its owner).
    auto leaf_acc = [&](uint64_t m, uint32_t leaf_node, int b) {
        return m & __builtin_amdgcn_ballot_w64(my_leaf[b] != leaf_node);
    };
    for (;;) {
        uint64_t om[kB];
        uint64_t any = 0;
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const uint64_t fm = __builtin_amdgcn_ballot_w64(probe(c, b));
            const uint64_t acc = mask[b] & fm & __builtin_amdgcn_ballot_w64(my_leaf[b] != node);
            const uint64_t open = mask[b] & ~fm;
            if (acc) pull(c, node, b, acc);
            tally(b, acc, open, 0ull);
            om[b] = open;
            any |= open;
        }
        // 1 when the walk goes on into a child; an opaque scalar integer
        // (a boolean crossing these blocks would become a lane mask)
        uint32_t go = 0u;
        if (any) {
            // the pair record: both children's records and indices, one
            // 64-byte line, 64-bit offsets (n < 2^30) (HIP vector types
            // cannot be loaded through the constant address space: the
            // record is read as dwords)
            const size_t rb = 4 * kPairStride * (size_t)node;
            const_as<float> *rec = pcm_s + rb;
            const float4 cl = make_float4(rec[0], rec[1], rec[2], rec[3]);
            const float4 cr = make_float4(rec[4], rec[5], rec[6], rec[7]);
            const uint2 ch = make_uint2(pcmu_s[rb + 8], pcmu_s[rb + 9]);
            {
                // one 64-byte load (the child indices are needed at once)
                PairPad pad;
                load_pad(pad, pcmu_s + rb);
                keep_live(pad);
            }
            const bool rleaf = ch.y >= leaf0, lleaf = ch.x >= leaf0;
            // the child walked next: right if internal, else left; chosen
            // here, beside the compare, so the choice stays a scalar select
            const uint32_t next = rleaf ? ch.x : ch.y;
            const float4 next_c = rleaf ? cl : cr;
            if (rleaf) {
#pragma unroll
                for (int b = 0; b < kB; ++b) {
                    const uint64_t acc = leaf_acc(om[b], ch.y, b);
                    if (acc) {
                        probe(cr, b);
                        pull(cr, ch.y, b, acc);
                    }
                    tally(b, acc, 0ull, 0ull);
                }
                if (lleaf) {
#pragma unroll
                    for (int b = 0; b < kB; ++b) {
                        const uint64_t acc = leaf_acc(om[b], ch.x, b);
                        if (acc) {
                            probe(cl, b);
                            pull(cl, ch.x, b, acc);
                        }
                        tally(b, acc, 0ull, 0ull);
                    }
                }
            }
            go = uniform_u32((rleaf && lleaf) ? 0u : 1u);
            if (!rleaf) {
                // push left, go right (the reference pops the right child first)
                if (sp < 64) {
                    st.template put_row<0>(sp, ch.x, om, cl);
                    ++sp;
                } else if (sp < stack_cap) {
                    st.template put_row<1>(sp & 63, ch.x, om, cl);
                    ++sp;
                } else {
                    // every lane stores the same flag: no divergent branch
                    // in the loop (it would make the node selects per-lane)
                    *overflow = 1u;
#pragma unroll
                    for (int b = 0; b < kB; ++b) tally(b, 0ull, 0ull, om[b]);
                }
            }
            if (go) {
                node = next;
                c = next_c;
#pragma unroll
                for (int b = 0; b < kB; ++b) mask[b] = om[b];
            }
        }
        if (!go) {
            if (sp == 0) break;
            --sp;
            st.get(sp >> 6, sp & 63, node, mask, c);
        }
    }
}

}  // namespace nbs
