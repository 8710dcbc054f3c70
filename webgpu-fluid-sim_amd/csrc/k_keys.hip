// k_keys.hip -- bounding-box reduction and Morton encoding.
//
// Reference: compute_morton_codes_main (src/shaders/compute/lbvh.wgsl:42-69)
// maps each body into a fixed [-1000,1000]^2 box and interleaves 16-bit
// coordinates.  NBODY_MORTON_BBOX30 instead quantises 15 bits per axis over
// the step's own square bounding box (origin at the box minimum, side = the
// larger extent), which keeps the keys meaningful for unit-scale inputs.
// Both are HBM-bound streams: 8 B of position in, 8 B of key + index out per
// body (12 B for the ensemble's 64-bit keys).  The arithmetic is compiled
// without contraction so every rounding matches the binary32 reference order.
#include "nbody_internal.h"

namespace nbs {

namespace {

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// grid = (max_chunks, num_sims); block 256.  Chunk c of sim s covers bodies
// [sim_off[s] + c*kBBoxChunk, ...).  fminf/fmaxf ignore NaN and are exact, so
// the reduction order does not change the result.
__global__ __launch_bounds__(256) void bbox_partial_kernel(const float2 *__restrict__ pos,
                                                           const int64_t *__restrict__ sim_off,
                                                           int max_chunks,
                                                           float4 *__restrict__ partial) {
    const int s = blockIdx.y;
    const int64_t lo = sim_off[s] + (int64_t)blockIdx.x * kBBoxChunk;
    const int64_t end = sim_off[s + 1];
    if (lo >= end) return;  // uniform: this sim has fewer chunks
    const int64_t hi = lo + kBBoxChunk < end ? lo + kBBoxChunk : end;
    float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
        float2 p = pos[i];
        mnx = fminf(mnx, p.x);
        mny = fminf(mny, p.y);
        mxx = fmaxf(mxx, p.x);
        mxy = fmaxf(mxy, p.y);
    }
    mnx = wave_min(mnx);
    mny = wave_min(mny);
    mxx = wave_max(mxx);
    mxy = wave_max(mxy);
    __shared__ float4 sh[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = make_float4(mnx, mny, mxx, mxy);
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 r = sh[0];
        for (int w = 1; w < 4; ++w) {
            r.x = fminf(r.x, sh[w].x);
            r.y = fminf(r.y, sh[w].y);
            r.z = fmaxf(r.z, sh[w].z);
            r.w = fmaxf(r.w, sh[w].w);
        }
        partial[(int64_t)s * max_chunks + blockIdx.x] = r;
    }
}

// grid = num_sims, block = one wave.
__global__ __launch_bounds__(64) void bbox_final_kernel(const int64_t *__restrict__ sim_off,
                                                        int max_chunks,
                                                        const float4 *__restrict__ partial,
                                                        float4 *__restrict__ bbox) {
    const int s = blockIdx.x;
    const int64_t ns = sim_off[s + 1] - sim_off[s];
    const int nch = (int)((ns + kBBoxChunk - 1) / kBBoxChunk);
    float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int c = threadIdx.x; c < nch; c += 64) {
        float4 r = partial[(int64_t)s * max_chunks + c];
        mnx = fminf(mnx, r.x);
        mny = fminf(mny, r.y);
        mxx = fmaxf(mxx, r.z);
        mxy = fmaxf(mxy, r.w);
    }
    mnx = wave_min(mnx);
    mny = wave_min(mny);
    mxx = wave_max(mxx);
    mxy = wave_max(mxy);
    if (threadIdx.x == 0) bbox[s] = make_float4(mnx, mny, mxx, mxy);
}

// Spread the low 16 bits of v to the even bit positions.
__device__ __forceinline__ uint32_t part1by1(uint32_t v) {
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// lbvh.wgsl:48-65 (fixed box, 16 bits per axis).
__device__ __forceinline__ uint32_t code_reference(float2 p) {
    const float scale = 2000.0f;
    const float half = 0.5f * scale;
    float ux = clamp01((p.x + half) / scale);
    float uy = clamp01((p.y + half) / scale);
    uint32_t x = (uint32_t)(ux * 65535.0f + 0.5f);
    uint32_t y = (uint32_t)(uy * 65535.0f + 0.5f);
    x = x > 65535u ? 65535u : x;
    y = y > 65535u ? 65535u : y;
    return part1by1(x) | (part1by1(y) << 1);
}

// 15 bits per axis over the square box at the bbox minimum.
__device__ __forceinline__ uint32_t code_bbox30(float2 p, float4 b) {
    float side = fmaxf(b.z - b.x, b.w - b.y);
    if (!(side > 0.0f)) side = 1.0f;
    float ux = clamp01((p.x - b.x) / side);
    float uy = clamp01((p.y - b.y) / side);
    uint32_t x = (uint32_t)(ux * 32767.0f + 0.5f);
    uint32_t y = (uint32_t)(uy * 32767.0f + 0.5f);
    x = x > 32767u ? 32767u : x;
    y = y > 32767u ? 32767u : y;
    return part1by1(x) | (part1by1(y) << 1);
}

template <int kMode, bool kEnsemble>
__global__ __launch_bounds__(256) void encode_kernel(const float2 *__restrict__ pos, int64_t n,
                                                     const float4 *__restrict__ bbox,
                                                     const uint8_t *__restrict__ body_sim,
                                                     uint32_t *__restrict__ keys32,
                                                     uint64_t *__restrict__ keys64,
                                                     uint32_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 p = pos[i];
    if constexpr (kEnsemble) {
        const uint32_t s = body_sim[i];
        keys64[i] = ((uint64_t)s << kSimShift) | code_bbox30(p, bbox[s]);
    } else if constexpr (kMode == 0) {
        keys32[i] = code_reference(p);
    } else {
        keys32[i] = code_bbox30(p, bbox[0]);
    }
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void clear_u32_kernel(uint32_t *__restrict__ p, int64_t words) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * 256)
        p[i] = 0u;
}

}  // namespace

void launch_bbox(const float2 *pos, const int64_t *sim_off, int num_sims, int max_chunks,
                 float4 *partial, float4 *bbox, hipStream_t s) {
    dim3 grid(max_chunks, num_sims);
    bbox_partial_kernel<<<grid, 256, 0, s>>>(pos, sim_off, max_chunks, partial);
    bbox_final_kernel<<<num_sims, 64, 0, s>>>(sim_off, max_chunks, partial, bbox);
}

void launch_encode(const float2 *pos, int64_t n, int morton_mode, const float4 *bbox,
                   const uint8_t *body_sim, uint32_t *keys32, uint64_t *keys64, uint32_t *idx,
                   hipStream_t s) {
    if (n <= 0) return;
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (body_sim)
        encode_kernel<1, true><<<grid, 256, 0, s>>>(pos, n, bbox, body_sim, keys32, keys64, idx);
    else if (morton_mode == 0)
        encode_kernel<0, false><<<grid, 256, 0, s>>>(pos, n, bbox, body_sim, keys32, keys64, idx);
    else
        encode_kernel<1, false><<<grid, 256, 0, s>>>(pos, n, bbox, body_sim, keys32, keys64, idx);
}

void launch_clear_u32(uint32_t *p, int64_t words, hipStream_t s) {
    if (words <= 0) return;
    int64_t blocks = (words + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    clear_u32_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, words);
}

}  // namespace nbs
