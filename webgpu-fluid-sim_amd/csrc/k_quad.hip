// k_quad.hip -- second mass moments of every node, for the quadrupole
// expansion of the walk (no reference counterpart: the reference's walk
// uses each node's mass and centre of mass only, barnes_hut.wgsl:92-97).
//
// A node's central second moments Q = sum_i m_i (x_i - c)(x_i - c)^T over
// its leaf range [lo, hi] (sorted positions) need no bottom-up pass:
//   * a node of at most kDirectMax leaves sums its range directly, in binary64,
//     twice (centre of mass, then the moments about it), from its leaf
//     records loaded at once into registers;
//   * a larger node takes differences of binary64 prefix sums of the leaf
//     moments (m, m x, m y, m x^2, m x y, m y^2) about the root's centre of
//     mass, Q = S2 - S1 S1^T / S0, where the cancellation costs at most
//     (total mass / node mass) * (extent / node size)^2 ~ (N / 8)^2 units of
//     binary64 rounding (about 1e-6 relative at 800k bodies, 2e-5 at 4M: a
//     correction to the quadrupole term, itself below theta^2 of the pull).
// The prefix sums are 64-byte records (six moments and two spare words), so
// a lookup is one line.
// Each internal node p stores its children's moments in the 24 padding bytes
// of its 64-byte pair record (pcm[4p+2].zw, pcm[4p+3]: left xx, xy, yy, then
// right xx, xy, yy), the line the pair walk loads anyway; the root's go to
// root[2].  Leaves have none (zeros).
//
// Kernels: quad_scan_block_kernel (inclusive prefix sums inside 1024-leaf
// blocks), quad_scan_top_kernel (one workgroup: exclusive prefix over the
// block totals), quad_nodes_kernel (one thread per internal node: both
// children's moments from three prefix lookups, or direct sums).
#include "nbody_internal.h"

namespace nbs {

namespace {

constexpr int kQuadBlock = 1024;
constexpr uint32_t kDirectMax = 8;
constexpr int kRec = 8;  // doubles per prefix record

struct Mom6 {
    double v[6];
};

__device__ __forceinline__ double shfl_up_d(double x, int d) {
    return __shfl_up(x, (unsigned)d, 64);
}

// Inclusive scan of 6 binary64 moments over the 1024 threads of a block.
__device__ __forceinline__ void block_scan6(Mom6 &a, Mom6 *wave_tot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double y = shfl_up_d(a.v[k], d);
            if (lane >= d) a.v[k] += y;
        }
    }
    if (lane == 63) wave_tot[wave] = a;
    __syncthreads();
    if (wave == 0 && lane < kQuadBlock / 64) {
        Mom6 w = wave_tot[lane];
#pragma unroll
        for (int d = 1; d < kQuadBlock / 64; d <<= 1) {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double y = shfl_up_d(w.v[k], d);
                if (lane >= d) w.v[k] += y;
            }
        }
        wave_tot[lane] = w;  // inclusive totals of waves 0..lane
    }
    __syncthreads();
    if (wave > 0) {
        const Mom6 off = wave_tot[wave - 1];
#pragma unroll
        for (int k = 0; k < 6; ++k) a.v[k] += off.v[k];
    }
}

__global__ __launch_bounds__(kQuadBlock) void quad_scan_block_kernel(
    const float4 *__restrict__ leaf, const float4 *__restrict__ root, uint32_t n,
    double *__restrict__ prefix, double *__restrict__ block_sum, uint32_t blocks) {
    __shared__ Mom6 wave_tot[kQuadBlock / 64];
    const uint32_t t = blockIdx.x * kQuadBlock + threadIdx.x;
    const double ox = root[0].x, oy = root[0].y;
    Mom6 a;
#pragma unroll
    for (int k = 0; k < 6; ++k) a.v[k] = 0.0;
    if (t < n) {
        const float4 L = leaf[t];
        const double m = L.z, dx = (double)L.x - ox, dy = (double)L.y - oy;
        a.v[0] = m;
        a.v[1] = m * dx;
        a.v[2] = m * dy;
        a.v[3] = m * dx * dx;
        a.v[4] = m * dx * dy;
        a.v[5] = m * dy * dy;
    }
    block_scan6(a, wave_tot);
    if (t < n) {
        double4 *r = reinterpret_cast<double4 *>(prefix + (size_t)kRec * t);
        r[0] = make_double4(a.v[0], a.v[1], a.v[2], a.v[3]);
        r[1] = make_double4(a.v[4], a.v[5], 0.0, 0.0);
    }
    if (threadIdx.x == kQuadBlock - 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) block_sum[(size_t)kRec * blockIdx.x + k] = a.v[k];
    }
}

// Exclusive prefix over the block totals, 1024 blocks per round with a carry.
// n == 1: the root is a leaf and has no moments.
__global__ __launch_bounds__(kQuadBlock) void quad_scan_top_kernel(double *__restrict__ block_sum,
                                                                   uint32_t blocks, uint32_t n,
                                                                   float4 *__restrict__ root) {
    __shared__ Mom6 wave_tot[kQuadBlock / 64];
    __shared__ Mom6 carry;
    if (n == 1 && threadIdx.x == 0) root[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) carry.v[k] = 0.0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < blocks; b0 += kQuadBlock) {
        const uint32_t b = b0 + threadIdx.x;
        Mom6 a, own;
#pragma unroll
        for (int k = 0; k < 6; ++k) own.v[k] = b < blocks ? block_sum[(size_t)kRec * b + k] : 0.0;
        a = own;
        block_scan6(a, wave_tot);
        const Mom6 c = carry;
        __syncthreads();  // every thread has read the carry
        if (b < blocks) {
#pragma unroll
            for (int k = 0; k < 6; ++k)
                block_sum[(size_t)kRec * b + k] = c.v[k] + (a.v[k] - own.v[k]);
        }
        if (threadIdx.x == kQuadBlock - 1) {
#pragma unroll
            for (int k = 0; k < 6; ++k) carry.v[k] = c.v[k] + a.v[k];
        }
        __syncthreads();
    }
}

// Inclusive prefix P(t) of the leaf moments (t = -1: zero).
__device__ __forceinline__ void prefix_at(const double *__restrict__ prefix,
                                          const double *__restrict__ block_off, int64_t t,
                                          double (&P)[6]) {
    if (t < 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) P[k] = 0.0;
        return;
    }
    const double4 *r = reinterpret_cast<const double4 *>(prefix + (size_t)kRec * t);
    const double4 *o = reinterpret_cast<const double4 *>(block_off + (size_t)kRec * (t >> 10));
    const double4 r0 = r[0], r1 = r[1], o0 = o[0], o1 = o[1];
    P[0] = r0.x + o0.x;
    P[1] = r0.y + o0.y;
    P[2] = r0.z + o0.z;
    P[3] = r0.w + o0.w;
    P[4] = r1.x + o1.x;
    P[5] = r1.y + o1.y;
}

// Central moments from the sums over a range (differences of prefixes).
__device__ __forceinline__ float3 moments_from(const double (&hi)[6], const double (&lo)[6]) {
    double S[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) S[k] = hi[k] - lo[k];
    if (!(S[0] > 0.0)) return make_float3(0.0f, 0.0f, 0.0f);
    return make_float3((float)(S[3] - S[1] * S[1] / S[0]), (float)(S[4] - S[1] * S[2] / S[0]),
                       (float)(S[5] - S[2] * S[2] / S[0]));
}

// Moments of a range of at most kDirectMax leaves, summed directly: the
// records are loaded together, then the centre of mass and the moments
// about it are summed in order (binary64).
__device__ __forceinline__ float3 direct_moments(const float4 *__restrict__ leaf, uint32_t lo,
                                                 uint32_t hi) {
    float4 L[kDirectMax];
#pragma unroll
    for (uint32_t k = 0; k < kDirectMax; ++k)
        L[k] = lo + k <= hi ? leaf[lo + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t cnt = hi - lo + 1u;
    double m = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < kDirectMax; ++k) {
        if (k < cnt) {
            m += (double)L[k].z;
            sx += (double)L[k].z * (double)L[k].x;
            sy += (double)L[k].z * (double)L[k].y;
        }
    }
    if (!(m > 0.0)) return make_float3(0.0f, 0.0f, 0.0f);
    const double cx = sx / m, cy = sy / m;
    double qxx = 0.0, qxy = 0.0, qyy = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < kDirectMax; ++k) {
        if (k < cnt) {
            const double dx = (double)L[k].x - cx, dy = (double)L[k].y - cy;
            qxx += (double)L[k].z * dx * dx;
            qxy += (double)L[k].z * dx * dy;
            qyy += (double)L[k].z * dy * dy;
        }
    }
    return make_float3((float)qxx, (float)qxy, (float)qyy);
}

__global__ __launch_bounds__(256) void quad_nodes_kernel(TreeBufs tb, uint32_t n,
                                                         const double *__restrict__ prefix,
                                                         const double *__restrict__ block_off) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nint = n - 1u;
    if (p >= nint) return;
    const uint2 r = tb.range[p];
    const uint2 ch = tb.child[p];
    const uint32_t g = ch.x < nint ? tb.range[ch.x].y : r.x;  // the left child's last leaf
    const bool big_l = ch.x < nint && g - r.x >= kDirectMax;
    const bool big_r = ch.y < nint && r.y - g > kDirectMax;
    double plo[6], pg[6], phi[6];
    if (big_l || big_r) prefix_at(prefix, block_off, g, pg);
    if (big_l) prefix_at(prefix, block_off, (int64_t)r.x - 1, plo);
    if (big_r) prefix_at(prefix, block_off, r.y, phi);
    float3 ql = make_float3(0.0f, 0.0f, 0.0f), qr = ql;
    if (ch.x < nint) ql = big_l ? moments_from(pg, plo) : direct_moments(tb.leaf, r.x, g);
    if (ch.y < nint) qr = big_r ? moments_from(phi, pg) : direct_moments(tb.leaf, g + 1u, r.y);
    float *rec = reinterpret_cast<float *>(tb.pcm + kPairStride * (size_t)p);
    *reinterpret_cast<float2 *>(rec + 10) = make_float2(ql.x, ql.y);
    *reinterpret_cast<float4 *>(rec + 12) = make_float4(ql.z, qr.x, qr.y, qr.z);
    if (p == 0u) {
        float3 q0;
        if (nint + 1u > kDirectMax) {
            double z[6], all[6];
            prefix_at(prefix, block_off, -1, z);
            prefix_at(prefix, block_off, nint, all);
            q0 = moments_from(all, z);
        } else {
            q0 = direct_moments(tb.leaf, 0u, nint);
        }
        tb.root[2] = make_float4(q0.x, q0.y, q0.z, 0.0f);
    }
}

}  // namespace

int64_t quad_prefix_words(int64_t cap) {
    const int64_t blocks = (cap + kQuadBlock - 1) / kQuadBlock;
    return kRec * (cap + blocks);
}

void launch_quad(const TreeBufs &tb, int64_t n, double *prefix, hipStream_t s) {
    if (n < 1) return;
    const uint32_t blocks = (uint32_t)((n + kQuadBlock - 1) / kQuadBlock);
    double *block_off = prefix + (size_t)kRec * n;
    quad_scan_block_kernel<<<blocks, kQuadBlock, 0, s>>>(tb.leaf, tb.root, (uint32_t)n, prefix,
                                                         block_off, blocks);
    quad_scan_top_kernel<<<1, kQuadBlock, 0, s>>>(block_off, blocks, (uint32_t)n, tb.root);
    if (n > 1)
        quad_nodes_kernel<<<(unsigned)((n - 1 + 255) / 256), 256, 0, s>>>(tb, (uint32_t)n, prefix,
                                                                         block_off);
}

}  // namespace nbs
