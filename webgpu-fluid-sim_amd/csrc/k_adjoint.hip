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

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
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
            inv[b] = inv_dist<kExact>(rx[b], ry[b], eps2);
            return passes(c.w, inv[b], theta);
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
