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
    __device__ __forceinline__ void get(int row, int l, uint32_t &nd, uint64_t (&m)[kB],
                                        float4 &c) const {
        const bool r = row != 0;
        nd = (uint32_t)__builtin_amdgcn_readlane(r ? node[1] : node[0], l);
        c.x = __int_as_float(__builtin_amdgcn_readlane(r ? cx[1] : cx[0], l));
        c.y = __int_as_float(__builtin_amdgcn_readlane(r ? cy[1] : cy[0], l));
        c.z = __int_as_float(__builtin_amdgcn_readlane(r ? cm[1] : cm[0], l));
        c.w = __int_as_float(__builtin_amdgcn_readlane(r ? cw[1] : cw[0], l));
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(r ? mlo[b][1] : mlo[b][0], l);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(r ? mhi[b][1] : mhi[b][0], l);
            m[b] = ((uint64_t)hi << 32) | lo;
        }
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
// entries already on `st`.  Each lane holds kB bodies (slot b: my_leaf[b],
// lane mask mask[b]); the wavefront's bodies share one traversal.
// visit(c, leaf, node, b, act, acc, open) is called by the whole wavefront
// for every visited node and slot (act: the slot's body evaluates it; the
// visitor sets acc/open).  Stops when the stack is empty.  A full stack
// drops both children like barnes_hut.wgsl:100-105, raises *overflow and
// calls drop(b, lane_dropped) per slot.  A wavefront visits a node at most
// once, so a tree from the build kernels ends within 2 * leaf0 + 1
// iterations.  There is no iteration counter: its compare-and-exit costs
// 12-19% of the walk time at 800k bodies.
template <int kB, typename Visit, typename Drop>
__device__ __forceinline__ void walk_tree(const float4 *pcm, const uint2 *child, uint32_t leaf0,
                                          const uint32_t (&my_leaf)[kB], int lane, uint32_t node,
                                          uint64_t (&mask)[kB], float4 c, RegStack<kB> &st,
                                          int sp, int stack_cap, uint32_t *overflow,
                                          Visit &&visit, Drop &&drop) {
    const_as<uint32_t> *child_s = (const_as<uint32_t> *)child;
    const_as<float> *pcmf_s = (const_as<float> *)pcm;
    for (;;) {
        const bool leaf = node >= leaf0;
        uint64_t om[kB];
        uint64_t any = 0;
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const bool act =
                __builtin_amdgcn_inverse_ballot_w64(mask[b]) && !(leaf && node == my_leaf[b]);
            bool acc, open;
            visit(c, leaf, node, b, act, acc, open);
            om[b] = __ballot(open);
            any |= om[b];
        }
        if (any) {
            // both children's records in one scalar load
            const uint32_t pn = node;
            const float4 cl = make_float4(pcmf_s[8 * pn], pcmf_s[8 * pn + 1], pcmf_s[8 * pn + 2],
                                          pcmf_s[8 * pn + 3]);
            const float4 cr = make_float4(pcmf_s[8 * pn + 4], pcmf_s[8 * pn + 5],
                                          pcmf_s[8 * pn + 6], pcmf_s[8 * pn + 7]);
            const uint32_t chl = child_s[2 * pn], chr = child_s[2 * pn + 1];
            if (chr >= leaf0) {
                // A leaf child has no subtree: its lanes take it now, at the
                // same place in the order.
#pragma unroll
                for (int b = 0; b < kB; ++b) {
                    bool a2, o2;
                    visit(cr, true, chr, b,
                          __builtin_amdgcn_inverse_ballot_w64(om[b]) && chr != my_leaf[b], a2, o2);
                }
                if (chl < leaf0) {
                    node = chl;
#pragma unroll
                    for (int b = 0; b < kB; ++b) mask[b] = om[b];
                    c = cl;
                    continue;
                }
#pragma unroll
                for (int b = 0; b < kB; ++b) {
                    bool a2, o2;
                    visit(cl, true, chl, b,
                          __builtin_amdgcn_inverse_ballot_w64(om[b]) && chl != my_leaf[b], a2, o2);
                }
                // both children done: fall through to the pop
            } else if (sp < stack_cap) {
                // push left, go right (the reference pops the right child first)
                st.put(sp >> 6, sp & 63, chl, om, cl);
                ++sp;
                node = chr;
#pragma unroll
                for (int b = 0; b < kB; ++b) mask[b] = om[b];
                c = cr;
                continue;
            } else {
                if (lane == 0) atomicOr(overflow, 1u);
#pragma unroll
                for (int b = 0; b < kB; ++b) drop(b, __builtin_amdgcn_inverse_ballot_w64(om[b]));
            }
        }
        if (sp == 0) break;
        --sp;
        st.get(sp >> 6, sp & 63, node, mask, c);
    }
}

}  // namespace nbs
