// walk_common.h -- the wavefront-coherent tree traversal shared by the
// forward walk (k_walk.hip) and the adjoint walk (k_adjoint.hip).
//
// The 64 lanes of a wavefront hold 64 consecutive Morton-ordered bodies and
// walk one depth-first traversal together.  A traversal entry is (node, lane
// mask, node record); every lane in the mask applies the opening test of its
// own body, and the lanes that open a node form the mask of its children.
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

// Register-resident stack of one wavefront: entry k lives in lane k & 63 of
// row k >> 6, one VGPR per field, so a push or pop needs no LDS access, no
// exec-mask change and no readfirstlane.  A push writes lane k of each field
// with v_writelane_b32 (the backend puts the lane select in M0 as GFX9
// requires), a pop reads it back with v_readlane_b32.  No memory is accessed.
struct RegStack {
    int node[2], mlo[2], mhi[2], cx[2], cy[2], cm[2], cw[2];

    __device__ __forceinline__ void put(int row, int lane_sel, uint32_t nd, uint64_t m,
                                        const float4 &c) {
        const int l = lane_sel;
        if (row == 0) {
            node[0] = nbs_writelane((int)nd, l, node[0]);
            mlo[0] = nbs_writelane((int)(uint32_t)m, l, mlo[0]);
            mhi[0] = nbs_writelane((int)(uint32_t)(m >> 32), l, mhi[0]);
            cx[0] = nbs_writelane(__float_as_int(c.x), l, cx[0]);
            cy[0] = nbs_writelane(__float_as_int(c.y), l, cy[0]);
            cm[0] = nbs_writelane(__float_as_int(c.z), l, cm[0]);
            cw[0] = nbs_writelane(__float_as_int(c.w), l, cw[0]);
        } else {
            node[1] = nbs_writelane((int)nd, l, node[1]);
            mlo[1] = nbs_writelane((int)(uint32_t)m, l, mlo[1]);
            mhi[1] = nbs_writelane((int)(uint32_t)(m >> 32), l, mhi[1]);
            cx[1] = nbs_writelane(__float_as_int(c.x), l, cx[1]);
            cy[1] = nbs_writelane(__float_as_int(c.y), l, cy[1]);
            cm[1] = nbs_writelane(__float_as_int(c.z), l, cm[1]);
            cw[1] = nbs_writelane(__float_as_int(c.w), l, cw[1]);
        }
    }
    __device__ __forceinline__ void get(int row, int lane_sel, uint32_t &nd, uint64_t &m,
                                        float4 &c) const {
        const int r = row;
        nd = (uint32_t)__builtin_amdgcn_readlane(r ? node[1] : node[0], lane_sel);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(r ? mlo[1] : mlo[0], lane_sel);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(r ? mhi[1] : mhi[0], lane_sel);
        m = ((uint64_t)hi << 32) | lo;
        c.x = __int_as_float(__builtin_amdgcn_readlane(r ? cx[1] : cx[0], lane_sel));
        c.y = __int_as_float(__builtin_amdgcn_readlane(r ? cy[1] : cy[0], lane_sel));
        c.z = __int_as_float(__builtin_amdgcn_readlane(r ? cm[1] : cm[0], lane_sel));
        c.w = __int_as_float(__builtin_amdgcn_readlane(r ? cw[1] : cw[0], lane_sel));
    }
};

// Opening test of one lane: lengths, distances and 1/sqrt exactly as the
// forward walk computes them (barnes_hut.wgsl:81-94), so every walk built on
// it takes the same decisions.  Returns inv = 1/sqrt(|r|^2 + eps^2).
template <bool kExact>
__device__ __forceinline__ float inv_dist(float rx, float ry, float eps2) {
    if constexpr (kExact) {
        const float d2 = rx * rx + ry * ry;
        return 1.0f / sqrtf(d2 + eps2);
    } else {
        return __builtin_amdgcn_rsqf(__builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, eps2)));
    }
}

// Depth-first traversal of one wavefront from (node, mask, c) with `sp`
// entries already on `st`.  visit(c, leaf, node, act, acc, open) is called
// by the whole wavefront for every visited node (act: this lane evaluates it;
// the visitor sets acc/open per lane).  Stops when the stack is empty.
// A full stack drops both children like barnes_hut.wgsl:100-105 and raises
// *overflow.  A wavefront visits a node at most once, so a valid tree ends
// within 2 * leaf0 + 1 iterations; the bound only guards against a corrupt
// tree and reports like an overflow.
template <bool kPrefetch, typename Visit>
__device__ __forceinline__ void walk_tree(const float4 *pcm, const uint2 *child, uint32_t leaf0,
                                          uint32_t my_leaf, int lane, uint32_t node,
                                          uint64_t mask, float4 c, RegStack &st, int sp,
                                          int stack_cap, uint32_t *overflow, Visit &&visit) {
    const_as<uint32_t> *child_s = (const_as<uint32_t> *)child;
    const_as<float> *pcmf_s = (const_as<float> *)pcm;
    for (uint32_t budget = 2u * leaf0 + 2u;; --budget) {
        if (budget == 0u) {
            if (lane == 0) atomicOr(overflow, 1u);
            break;
        }
        const bool leaf = node >= leaf0;
        // Both children's records come in one scalar load (used if a lane
        // opens the node).  With kPrefetch the load is issued before the test
        // for every internal node so its latency overlaps the lane arithmetic;
        // a leaf reads pair 0, which is allocated for every n.
        const uint32_t pn = (kPrefetch && leaf) ? 0u : node;
        float4 cl, cr;
        uint32_t chl = 0, chr = 0;
        if constexpr (kPrefetch) {
            cl = make_float4(pcmf_s[8 * pn], pcmf_s[8 * pn + 1], pcmf_s[8 * pn + 2],
                             pcmf_s[8 * pn + 3]);
            cr = make_float4(pcmf_s[8 * pn + 4], pcmf_s[8 * pn + 5], pcmf_s[8 * pn + 6],
                             pcmf_s[8 * pn + 7]);
            chl = child_s[2 * pn];
            chr = child_s[2 * pn + 1];
        }
        const bool act = __builtin_amdgcn_inverse_ballot_w64(mask) && !(leaf && node == my_leaf);
        bool acc, open;
        visit(c, leaf, node, act, acc, open);
        if constexpr (kPrefetch) {
            // keeps the loads above the test: without a use here the compiler
            // sinks them into the branch below, after the vote
            asm volatile("" ::"s"(cl.x), "s"(cl.w), "s"(cr.x), "s"(cr.w), "s"(chl), "s"(chr));
        }
        const uint64_t om = __ballot(open);
        if (om) {
            if constexpr (!kPrefetch) {
                cl = make_float4(pcmf_s[8 * pn], pcmf_s[8 * pn + 1], pcmf_s[8 * pn + 2],
                                 pcmf_s[8 * pn + 3]);
                cr = make_float4(pcmf_s[8 * pn + 4], pcmf_s[8 * pn + 5], pcmf_s[8 * pn + 6],
                                 pcmf_s[8 * pn + 7]);
                chl = child_s[2 * pn];
                chr = child_s[2 * pn + 1];
            }
            const bool in = __builtin_amdgcn_inverse_ballot_w64(om);
            if (chr >= leaf0) {
                // A leaf child has no subtree: its lanes take it now, at the
                // same place in the order.
                bool a2, o2;
                visit(cr, true, chr, in && chr != my_leaf, a2, o2);
                if (chl < leaf0) {
                    node = chl;
                    mask = om;
                    c = cl;
                    continue;
                }
                visit(cl, true, chl, in && chl != my_leaf, a2, o2);
                // both children done: fall through to the pop
            } else if (sp < stack_cap) {
                // push left, go right (the reference pops the right child first)
                st.put(sp >> 6, sp & 63, chl, om, cl);
                ++sp;
                node = chr;
                mask = om;
                c = cr;
                continue;
            } else if (lane == 0) {
                atomicOr(overflow, 1u);
            }
        }
        if (sp == 0) break;
        --sp;
        st.get(sp >> 6, sp & 63, node, mask, c);
    }
}

}  // namespace nbs
