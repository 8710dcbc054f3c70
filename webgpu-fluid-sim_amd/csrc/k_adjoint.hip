// k_adjoint.hip -- reverse-mode (adjoint) kernels of the Barnes-Hut step.
//
// No reference counterpart: the reference only steps forward
// (src/simulation.ts:386-425).  The adjoint differentiates the step with the
// tree topology and every opening decision frozen at their forward values:
// the tree is rebuilt from the checkpointed positions of the step (the
// pipeline is deterministic, so it is the same tree) and walked with the
// forward walk's own traversal and test.
//
// Forward step (constant softening):  a_i = sum_n G M_n r_in s^-3/2,
//   r_in = c_n - x_i, s = |r|^2 + eps^2;  v' = v + a dt;  x' = x + v' dt.
// Reverse step, given (lx, lv) = dL/d(x', v'):
//   lv <- lv + lx dt;  la = lv dt                        (adjoint_pre_kernel)
//   per accepted (i, n), g = la_i:
//     h = G M_n (s^-3/2 g - 3 s^-5/2 (r.g) r)
//     lx_i -= h;  lc_n += h;  lM_n += G s^-3/2 (r.g)    (adjoint_walk_kernel)
//   node cotangents to leaves through c_n = sum m_j x_j / M_n, M_n = sum m_j:
//     lx_j += m_j sum_{n ∋ j} lc_n / M_n
//     lm_j += sum_{n ∋ j} [ x_j . lc_n / M_n - c_n . lc_n / M_n + lM_n ]
//                                                        (adjoint_fold_kernel)
#include "walk_common.h"

namespace nbs {

namespace {

// Sum over the wavefront (all lanes active): DPP adds inside each row of 16
// lanes (lane ^ 1, lane ^ 2, the other quad, the other half-row), then the
// four row sums, ((r0 + r1) + (r2 + r3)).  A butterfly of __shfl_xor goes
// through the LDS crossbar (ds_bpermute): six of them per sum.
template <int kCtrl>
__device__ __forceinline__ float dpp_f32(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f32<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f32<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f32<0x141>(v);  // row_half_mirror
    v += dpp_f32<0x140>(v);  // row_mirror
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (r0 + r1) + (r2 + r3);
}

__global__ __launch_bounds__(256) void adjoint_pre_kernel(const float2 *__restrict__ lx,
                                                          float2 *__restrict__ lv,
                                                          float2 *__restrict__ la, uint32_t n,
                                                          float dt) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 x = lx[i];
    float2 v = lv[i];
    v.x = __builtin_fmaf(x.x, dt, v.x);
    v.y = __builtin_fmaf(x.y, dt, v.y);
    lv[i] = v;
    la[i] = make_float2(v.x * dt, v.y * dt);
}

// The forward walk's traversal and opening test (walk_common.h), with the
// VJP of the pull on accepted nodes instead of the pull; kB bodies per lane
// as in the forward walk.
template <bool kExact, int kB>
__global__ __launch_bounds__(256) void adjoint_walk_kernel(const AdjointArgs a, int stack_cap) {
    const int lane = threadIdx.x & 63;
    const uint32_t n = a.n;
    const uint32_t first = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * (64u * kB);
    const uint32_t leaf0 = n - 1u;

    uint32_t t[kB], my_leaf[kB], bi[kB];
    bool valid[kB];
    float px[kB], py[kB], gx[kB], gy[kB], hx_sum[kB], hy_sum[kB];
    uint64_t mask[kB];
    uint64_t live = 0;
#pragma unroll
    for (int b = 0; b < kB; ++b) {
        t[b] = first + 64u * b + lane;
        valid[b] = t[b] < n;
        my_leaf[b] = leaf0 + t[b];
        px[b] = py[b] = gx[b] = gy[b] = hx_sum[b] = hy_sum[b] = 0.0f;
        bi[b] = 0;
        if (valid[b]) {
            const float4 me = a.leaf[t[b]];
            px[b] = me.x;
            py[b] = me.y;
            bi[b] = a.idx[t[b]];
            const float2 g = a.la[bi[b]];
            gx[b] = g.x;
            gy[b] = g.y;
        }
        mask[b] = __ballot(valid[b]);
        live |= mask[b];
    }
    if (live == 0) return;

    const float G = a.grav, theta = a.theta, eps2 = a.eps * a.eps;
    RegStack<kB> st;
    float rx[kB], ry[kB], inv[kB];
    walk_tree<kB>(
        a.pcm, leaf0, my_leaf, 0u, mask, a.root[0], st, 0, stack_cap,
        a.overflow,
        [&](const float4 &c, int b) {
            rx[b] = c.x - px[b];
            ry[b] = c.y - py[b];
            return probe_node<kExact>(c, rx[b], ry[b], eps2, theta, a.open_scale, a.open_bias,
                                      inv[b]);
        },
        [&](const float4 &c, uint32_t node, int b, uint64_t acc_mask) {
            const bool acc = __builtin_amdgcn_inverse_ballot_w64(acc_mask);
            const float inv2 = inv[b] * inv[b];
            const float inv3 = inv2 * inv[b];
            const float rg = rx[b] * gx[b] + ry[b] * gy[b];
            const float gm = G * c.z;
            const float k3 = 3.0f * inv2 * rg;
            const float hx = acc ? gm * inv3 * (gx[b] - k3 * rx[b]) : 0.0f;
            const float hy = acc ? gm * inv3 * (gy[b] - k3 * ry[b]) : 0.0f;
            const float dm = acc ? G * inv3 * rg : 0.0f;
            hx_sum[b] += hx;
            hy_sum[b] += hy;
            // node cotangent: one atomic per wavefront, slot and component
            const float sx = wave_sum(hx), sy = wave_sum(hy), sm = wave_sum(dm);
            if (lane == 0) {
                float4 *cot = a.cot + node;
                atomicAdd(&cot->x, sx);
                atomicAdd(&cot->y, sy);
                atomicAdd(&cot->z, sm);
            }
        },
        [](int, uint64_t, uint64_t, uint64_t) {});
#pragma unroll
    for (int b = 0; b < kB; ++b) {
        if (!valid[b]) continue;
        float2 l = a.lx[bi[b]];
        l.x -= hx_sum[b];
        l.y -= hy_sum[b];
        a.lx[bi[b]] = l;
    }
}

// FAST mode: the pair walk's traversal (k_walk.hip: pair_walk_chunk) —
// one iteration per opened node, both children tested from one pair record,
// (node, mask) stack in three VGPRs — with the VJP of the pull for the lanes
// that accept a child.  The own leaf is left out by a lane compare (its
// forward term is an exact zero, its VJP terms would cancel only in the fold).
// The decisions are the forward pair walk's, so the node cotangents and body
// terms are those of the forward sum it differentiates.
//
// Node cotangents are first summed per workgroup in an LDS hash table (the
// workgroup's Morton-adjacent wavefronts accept many of the same nodes) and
// flushed to memory with one atomic per node and component at the end; a
// crowded table sends the rest straight to memory.
constexpr int kCotBits = 11;
constexpr uint32_t kCotSlots = 1u << kCotBits;
constexpr int kCotProbes = 8;
template <bool kOff32>
__global__ __launch_bounds__(256) void adjoint_pair_kernel(const AdjointArgs a) {
    __shared__ uint32_t s_node[kCotSlots];
    __shared__ float s_cx[kCotSlots], s_cy[kCotSlots], s_cm[kCotSlots];
    for (uint32_t k = threadIdx.x; k < kCotSlots; k += blockDim.x) {
        s_node[k] = kNoNode;
        s_cx[k] = s_cy[k] = s_cm[k] = 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint32_t n = a.n;
    const uint32_t t = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64u + (uint32_t)lane;
    const bool valid = t < n;
    const uint32_t leaf0 = n - 1u;
    const uint32_t my_leaf = leaf0 + t;
    float px = 0.0f, py = 0.0f, gx = 0.0f, gy = 0.0f, hx_sum = 0.0f, hy_sum = 0.0f;
    uint32_t bi = 0;
    float2 lx0 = make_float2(0.0f, 0.0f);  // the epilogue's input, read now
    if (valid) {
        const float4 me = a.leaf[t];
        px = me.x;
        py = me.y;
        bi = a.idx[t];
        const float2 g = a.la[bi];
        lx0 = a.lx[bi];
        gx = g.x;
        gy = g.y;
    }
    const uint64_t live = __ballot(valid);
    const float G = a.grav, theta = a.theta, eps2 = a.eps * a.eps;
    auto probe = [&](const float4 &c, float &rx, float &ry, float &inv) {
        rx = c.x - px;
        ry = c.y - py;
        return probe_node<false>(c, rx, ry, eps2, theta, a.open_scale, a.open_bias, inv);
    };
    // The VJP of the pull of a node (record c) for the lanes of acc: the
    // lane's terms (hx, hy, dm), added to its own sums by the caller.
    auto vjp_terms = [&](const float4 &c, float rx, float ry, float inv, uint64_t acc, float &hx,
                         float &hy, float &dm) {
        const bool on = __builtin_amdgcn_inverse_ballot_w64(acc);
        const float inv2 = inv * inv;
        const float inv3 = inv2 * inv;
        const float rg = rx * gx + ry * gy;
        const float gm = G * c.z;
        const float k3 = 3.0f * inv2 * rg;
        hx = on ? gm * inv3 * (gx - k3 * rx) : 0.0f;
        hy = on ? gm * inv3 * (gy - k3 * ry) : 0.0f;
        dm = on ? G * inv3 * rg : 0.0f;
    };
    // node cotangent (one call per wavefront, node): into the workgroup's
    // table, or to memory when its probe run is full
    auto cot_add = [&](uint32_t node, float sx, float sy, float sm) {
        uint32_t h = (node * 2654435761u) >> (32 - kCotBits);
        for (int p = 0; p < kCotProbes; ++p) {
            const uint32_t k = atomicCAS(&s_node[h], kNoNode, node);
            if (k == kNoNode || k == node) {
                atomicAdd(&s_cx[h], sx);
                atomicAdd(&s_cy[h], sy);
                atomicAdd(&s_cm[h], sm);
                return;
            }
            h = (h + 1u) & (kCotSlots - 1u);
        }
        float4 *cot = a.cot + node;
        atomicAdd(&cot->x, sx);
        atomicAdd(&cot->y, sy);
        atomicAdd(&cot->z, sm);
    };
    auto vjp = [&](const float4 &c, uint32_t node, float rx, float ry, float inv, uint64_t acc) {
        float hx, hy, dm;
        vjp_terms(c, rx, ry, inv, acc, hx, hy, dm);
        hx_sum += hx;
        hy_sum += hy;
        const float sx = wave_sum(hx), sy = wave_sum(hy), sm = wave_sum(dm);
        if (lane == 0) cot_add(node, sx, sy, sm);
    };
    if (live) {
        PairStack st;
        st.put(0, 0u, 0ull);  // the sentinel
        int sp = 1;
        {
            // the root, tested on its own record
            const float4 c = a.root[0];
            float rx, ry, inv;
            const uint64_t pass = __ballot(probe(c, rx, ry, inv));
            uint64_t acc = live & pass;
            if (leaf0 == 0u) acc &= __ballot(my_leaf != 0u);  // a one-body tree
            const uint64_t open = live & ~pass;
            if (acc) vjp(c, 0u, rx, ry, inv, acc);
            if (open) {
                st.put(sp, 0u, open);
                ++sp;
            }
        }
        if (sp > 1) {
            const_as<float> *pcm_s = (const_as<float> *)a.pcm;
            const_as<uint32_t> *pcmu_s = (const_as<uint32_t> *)a.pcm;
            uint32_t node;
            uint64_t mask;
            --sp;
            st.get(sp, node, mask);
            const size_t rb = 4 * kPairStride * (size_t)node;
            float4 cl = make_float4(uniform_f32(pcm_s[rb + 0]), uniform_f32(pcm_s[rb + 1]),
                                    uniform_f32(pcm_s[rb + 2]), uniform_f32(pcm_s[rb + 3]));
            float4 cr = make_float4(uniform_f32(pcm_s[rb + 4]), uniform_f32(pcm_s[rb + 5]),
                                    uniform_f32(pcm_s[rb + 6]), uniform_f32(pcm_s[rb + 7]));
            uint2 ch = make_uint2(uniform_u32(pcmu_s[rb + 8]), uniform_u32(pcmu_s[rb + 9]));
            uint32_t n_over = 0;
            PairPad pad = {};
            for (;;) {
                keep_live(pad);  // one 64-byte load per record
                float rxl, ryl, invl, rxr, ryr, invr;
                const uint64_t bl = __ballot(probe(cl, rxl, ryl, invl));
                const uint64_t br = __ballot(probe(cr, rxr, ryr, invr));
                const uint64_t accl = mask & bl & __ballot(my_leaf != ch.x);
                const uint64_t accr = mask & br & __ballot(my_leaf != ch.y);
                const uint64_t openl = mask & ~bl, openr = mask & ~br;
                const uint32_t hl = nonzero_flag(openl), hr = nonzero_flag(openr);
                st.put(sp, ch.x, openl);
                const uint32_t top = (uint32_t)sp + (hl & hr);
                n_over += top >> 6;  // top == 64: the left child is dropped
                const uint32_t any = hl | hr;
                uint32_t next = hr ? ch.y : ch.x;
                uint64_t nmask = hr ? openr : openl;
                sp = any ? (int)min(top, (uint32_t)kPairStackTop) : sp - 1;
                if (!any) st.get(sp, next, nmask);
                // the next record's loads (the sentinel's when the walk is
                // done), issued before this node's pulls
                const float4 pl = cl, pr = cr;
                const uint2 pch = ch;
                {
                    const_as<float> *rf;
                    if constexpr (kOff32)
                        rf = (const_as<float> *)((const_as<char> *)pcm_s + (uint32_t)(next * 64u));
                    else
                        rf = pcm_s + 4 * kPairStride * (size_t)next;
                    const_as<uint32_t> *ru = (const_as<uint32_t> *)rf;
                    cl = make_float4(rf[0], rf[1], rf[2], rf[3]);
                    cr = make_float4(rf[4], rf[5], rf[6], rf[7]);
                    ch = make_uint2(ru[8], ru[9]);
                    load_pad(pad, ru);
                }
                if (accl | accr) {
                    // both children's terms and their six wavefront sums together
                    float hxr, hyr, dmr, hxl, hyl, dml;
                    vjp_terms(pr, rxr, ryr, invr, accr, hxr, hyr, dmr);
                    vjp_terms(pl, rxl, ryl, invl, accl, hxl, hyl, dml);
                    hx_sum += hxr;
                    hy_sum += hyr;
                    hx_sum += hxl;
                    hy_sum += hyl;
                    const float sxr = wave_sum(hxr), syr = wave_sum(hyr), smr = wave_sum(dmr);
                    const float sxl = wave_sum(hxl), syl = wave_sum(hyl), sml = wave_sum(dml);
                    if (lane == 0) {
                        if (accr) cot_add(pch.y, sxr, syr, smr);
                        if (accl) cot_add(pch.x, sxl, syl, sml);
                    }
                }
                mask = nmask;
                if (mask == 0) break;
            }
            if (n_over) *a.overflow = 1u;
        }
    }
    if (valid) {
        float2 l = lx0;
        l.x -= hx_sum;
        l.y -= hy_sum;
        a.lx[bi] = l;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kCotSlots; k += blockDim.x) {
        const uint32_t node = s_node[k];
        if (node == kNoNode) continue;
        float4 *cot = a.cot + node;
        atomicAdd(&cot->x, s_cx[k]);
        atomicAdd(&cot->y, s_cy[k]);
        atomicAdd(&cot->z, s_cm[k]);
    }
}

// Record {com, mass, length} of a node: root record, leaf record, or the
// child slot in its parent's pair.
__device__ __forceinline__ float4 node_data(const AdjointArgs &a, uint32_t node) {
    if (node == 0u) return a.root[0];
    const uint32_t pw = a.parent[node];
    return a.pcm[pcm_slot(pw)];
}

// One thread per sorted leaf: sum lc_n / M_n and lM_n - c_n . lc_n / M_n over
// the leaf and its ancestors, then apply to the body (see the header).
__global__ __launch_bounds__(256) void adjoint_fold_kernel(const AdjointArgs a) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = a.n;
    if (t >= n) return;
    const float4 me = a.leaf[t];
    uint32_t cur = n - 1u + t;
    float pax = 0.0f, pay = 0.0f, pb = 0.0f;
    for (int level = 0; level < 128; ++level) {
        const float4 c = (cur == n - 1u + t) ? me : node_data(a, cur);
        const float4 l = a.cot[cur];
        if (c.z > 0.0f) {
            const float ax = l.x / c.z, ay = l.y / c.z;
            pax += ax;
            pay += ay;
            pb += l.z - (c.x * ax + c.y * ay);
        }
        if (cur == 0u) break;
        const uint32_t par = a.parent[cur] & kParentMask;
        if (par >= n - 1u) break;
        cur = par;
    }
    const uint32_t bi = a.idx[t];
    float2 lx = a.lx[bi];
    lx.x += me.z * pax;
    lx.y += me.z * pay;
    a.lx[bi] = lx;
    a.lm[bi] += me.x * pax + me.y * pay + pb;
}

template <bool kExact, int kB>
void launch_adjoint_walk_t(const AdjointArgs &a, int cap, hipStream_t s) {
    const unsigned block = 256, per_block = block * kB;
    const unsigned grid = (a.n + per_block - 1) / per_block;
    adjoint_walk_kernel<kExact, kB><<<grid, block, 0, s>>>(a, cap);
}

}  // namespace

void launch_adjoint_pre(const float2 *lx, float2 *lv, float2 *la, uint32_t n, float dt,
                        hipStream_t s) {
    if (n == 0) return;
    adjoint_pre_kernel<<<(n + 255) / 256, 256, 0, s>>>(lx, lv, la, n, dt);
}

void launch_adjoint_walk(const AdjointArgs &a, bool exact, int stack, int lane_bodies,
                         hipStream_t s) {
    if (a.n == 0) return;
    if (!exact) {
        const unsigned grid = (a.n + 255u) / 256u;
        if (a.n <= (1u << 26))
            adjoint_pair_kernel<true><<<grid, 256, 0, s>>>(a);
        else
            adjoint_pair_kernel<false><<<grid, 256, 0, s>>>(a);
        return;
    }
    const int cap = stack < 64 ? 64 : (stack > 128 ? 128 : stack);
    if (lane_bodies == 2) {
        if (exact) launch_adjoint_walk_t<true, 2>(a, cap, s);
        else launch_adjoint_walk_t<false, 2>(a, cap, s);
    } else {
        if (exact) launch_adjoint_walk_t<true, 1>(a, cap, s);
        else launch_adjoint_walk_t<false, 1>(a, cap, s);
    }
}

void launch_adjoint_fold(const AdjointArgs &a, hipStream_t s) {
    if (a.n == 0) return;
    adjoint_fold_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
}

}  // namespace nbs
