// k_keys.hip -- bounding-box reduction and Morton encoding.
//
// Reference: compute_morton_codes_main (src/shaders/compute/lbvh.wgsl:42-69)
// maps each body into a fixed [-1000,1000]^2 box and interleaves 16-bit
// coordinates.  NBODY_MORTON_BBOX30 instead quantises 15 bits per axis over
// the step's own square bounding box (origin at the box minimum, side = the
// larger extent), which keeps the keys meaningful for unit-scale inputs.
// Both are HBM-bound streams: 8 B of position in, 8 B of key + index out per
// body (12 B for the ensemble's 64-bit keys).  The encode also counts the
// radix sort's digits (every pass) for the keys it writes.  The arithmetic is
// compiled without contraction so every rounding matches the binary32
// reference order.
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

// grid = num_sims, block = one wave.  slots (one sim): the walk's
// bounding-box accumulators are set to this box (slot 0) and empty slots.
__global__ __launch_bounds__(64) void bbox_final_kernel(const int64_t *__restrict__ sim_off,
                                                        int max_chunks,
                                                        const float4 *__restrict__ partial,
                                                        float4 *__restrict__ bbox,
                                                        uint32_t *__restrict__ slots) {
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
    if (slots) {
        // (a box without bodies, +inf/-inf, stays an empty slot)
        const bool first = threadIdx.x == 0;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (first && mnx <= mxx) {
            v = make_uint4(~float_order(mnx), ~float_order(mny), float_order(mxx),
                           float_order(mxy));
        }
        for (int k = threadIdx.x; k < kBBoxSlots; k += 64)
            reinterpret_cast<uint4 *>(slots)[k] = k == 0 ? v : make_uint4(0u, 0u, 0u, 0u);
    }
}

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

// The bounding box held by the walk's accumulators (all 64 lanes of the
// wave take part; kBBoxSlots / 64 slots per lane).
__device__ __forceinline__ float4 bbox_from_slots(const uint32_t *__restrict__ slots) {
    static_assert(kBBoxSlots % 64 == 0, "slots per lane");
    uint4 u[kBBoxSlots / 64];  // all loads in flight before the first max
#pragma unroll
    for (int j = 0; j < kBBoxSlots / 64; ++j)
        u[j] = reinterpret_cast<const uint4 *>(slots)[j * 64 + (threadIdx.x & 63)];
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int j = 0; j < kBBoxSlots / 64; ++j)
        v = make_uint4(max(v.x, u[j].x), max(v.y, u[j].y), max(v.z, u[j].z), max(v.w, u[j].w));
    const uint32_t mnx = wave_max_u32(v.x), mny = wave_max_u32(v.y);
    const uint32_t mxx = wave_max_u32(v.z), mxy = wave_max_u32(v.w);
    // an empty accumulator reads as the identity box (+inf, -inf)
    return make_float4(mnx ? float_unorder(~mnx) : INFINITY, mny ? float_unorder(~mny) : INFINITY,
                       mxx ? float_unorder(mxx) : -INFINITY, mxy ? float_unorder(mxy) : -INFINITY);
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

// Keys of bodies [blockIdx.x * kEncodeTile, +kEncodeTile), 4 per thread
// (item it of thread t is body base + it * 1024 + t: coalesced), and the
// digit counts of every sort pass for them: LDS histograms, then one global
// atomic per non-empty bin per workgroup (large workgroups keep those few).
// The sort needs these counts before its first pass; counting here saves a
// kernel and a read of the keys.
constexpr int kEncodeThreads = 1024;
constexpr int kEncodeTile = 4 * kEncodeThreads;
template <int kMode, bool kEnsemble, int RB>
__global__ __launch_bounds__(kEncodeThreads) void encode_kernel(const float2 *__restrict__ pos, int64_t n,
                                                     float4 *__restrict__ bbox,
                                                     const uint32_t *__restrict__ slots,
                                                     const uint8_t *__restrict__ body_sim,
                                                     uint32_t *__restrict__ keys32,
                                                     uint64_t *__restrict__ keys64,
                                                     uint32_t *__restrict__ idx,
                                                     uint32_t *__restrict__ hist, int passes) {
    constexpr int D = 1 << RB;
    constexpr uint32_t kMask = (uint32_t)D - 1u;
    __shared__ uint32_t sh[kSortMaxPasses * D];
    const int bins = passes * D;
    constexpr int kItems = kEncodeTile / kEncodeThreads;
    const int64_t base = (int64_t)blockIdx.x * kEncodeTile;
    // the positions are loaded first, so their latency overlaps the box's
    // (the slots come from the walk's device-scope atomics, past the L2)
    float2 p[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t i = base + it * kEncodeThreads + threadIdx.x;
        p[it] = i < n ? pos[i] : make_float2(0.0f, 0.0f);
    }
    for (int i = threadIdx.x; i < bins; i += kEncodeThreads) sh[i] = 0u;
    float4 box = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (kMode == 1 && !kEnsemble) {
        if (slots) {
            box = bbox_from_slots(slots);  // every wave reduces the 64 slots itself
            if (blockIdx.x == 0 && threadIdx.x == 0) bbox[0] = box;
        } else {
            box = bbox[0];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t i = base + it * kEncodeThreads + threadIdx.x;
        if (i >= n) break;
        uint64_t key;
        if constexpr (kEnsemble) {
            const uint32_t s = body_sim[i];
            key = ((uint64_t)s << kSimShift) | code_bbox30(p[it], bbox[s]);
            keys64[i] = key;
        } else if constexpr (kMode == 0) {
            key = code_reference(p[it]);
            keys32[i] = (uint32_t)key;
        } else {
            key = code_bbox30(p[it], box);
            keys32[i] = (uint32_t)key;
        }
        idx[i] = (uint32_t)i;
        for (int q = 0; q < passes; ++q)
            atomicAdd(&sh[q * D + (uint32_t)((key >> (RB * q)) & kMask)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < bins; i += kEncodeThreads) {
        const uint32_t c = sh[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// The bucket sort's encode and placing in one pass: the keys, each body's
// bucket (the number of splitters <= its key: a binary search in LDS), then
// the tile's bodies grouped by bucket, stably (a
// body's place among its bucket's bodies of the tile is its rank by input
// order: ballots over the bucket id's bits, as the radix passes rank a
// digit), at the tile's own positions.  The tile writes, for every bucket b,
// its count and first position among the tile's bodies to tcount[b * tiles +
// tile] and tstart[b * tiles + tile], and adds its counts to hist; the
// per-bucket sort gathers bucket b's bodies tile after tile, which is input
// order, with no pass over all bodies between the two kernels.  Item it of
// lane l in wave w is tile position w * 64 * kItems + it * 64 + l, so that
// (w, it, l) order is input order.
template <int kMode, int RB, int kItems>
__global__ __launch_bounds__(kEncodeThreads) void encode_group_kernel(
    const float2 *__restrict__ pos, int64_t n, float4 *__restrict__ bbox,
    const uint32_t *__restrict__ slots, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
    const uint32_t *__restrict__ split, int nb, uint32_t *__restrict__ hist,
    uint32_t *__restrict__ tcount, uint32_t *__restrict__ tstart, int tiles) {
    constexpr int D = 1 << RB;
    constexpr int kWaves = kEncodeThreads / kWave;
    constexpr int kTile = kItems * kEncodeThreads;
    static_assert(D <= kEncodeThreads && D >= kMaxBuckets / 4, "one bucket per owner thread");
    __shared__ uint32_t s_split[kMaxBuckets];
    __shared__ uint32_t s_wcnt[kWaves * D];  // wave bucket counts, then wave offsets
    __shared__ uint32_t s_lstart[D];         // bucket start inside the tile
    __shared__ uint32_t s_shw[kWaves];
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int tile = (int)blockIdx.x;
    const int64_t base = (int64_t)tile * kTile;
    float2 p[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t i = base + wid * (64 * kItems) + it * 64 + lane;
        p[it] = i < n ? pos[i] : make_float2(0.0f, 0.0f);
    }
    // the splitters (one per thread: nb - 1 < kMaxBuckets <= kEncodeThreads)
    // are read into registers and written to LDS after the box's loads are
    // issued: one memory round trip for positions, splitters and box
    static_assert(kMaxBuckets <= kEncodeThreads, "one splitter per thread");
    const bool has_split = tid < nb - 1;
    const uint32_t sv = has_split ? split[tid] : 0u;
    for (int j = tid; j < kWaves * D; j += kEncodeThreads) s_wcnt[j] = 0u;
    float4 box = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (kMode == 1) {
        if (slots) {
            box = bbox_from_slots(slots);
            if (tile == 0 && tid == 0) bbox[0] = box;
        } else {
            box = bbox[0];
        }
    }
    if (has_split) s_split[tid] = sv;
    __syncthreads();
    const int ns = nb - 1;
    int top = 1;
    while (top * 2 <= ns) top *= 2;
    uint32_t key[kItems], dg[kItems], rank[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t i = base + wid * (64 * kItems) + it * 64 + lane;
        key[it] = 0u;
        dg[it] = (uint32_t)D;  // no body
        if (i < n) {
            key[it] = kMode == 0 ? code_reference(p[it]) : code_bbox30(p[it], box);
            int b = 0;  // splitters <= key
            if (ns > 0)
                for (int step = top; step >= 1; step >>= 1)
                    if (b + step <= ns && s_split[b + step - 1] <= key[it]) b += step;
            dg[it] = (uint32_t)b;
        }
    }
    // stable rank inside the wave among the lanes of the same bucket
    uint32_t *wc = s_wcnt + wid * D;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t d = dg[it];
        const bool ok = d < (uint32_t)D;
        const uint64_t valid = __ballot(ok);
        uint32_t plo = (uint32_t)valid, phi = (uint32_t)(valid >> 32);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const uint32_t m = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
            const uint64_t bb = __ballot(m != 0u);
            plo &= ~((uint32_t)bb ^ m);
            phi &= ~((uint32_t)(bb >> 32) ^ m);
        }
        const uint64_t peers = ((uint64_t)phi << 32) | plo;
        const uint32_t below = lanes_below(peers);
        const uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t before = 0;
        if (ok) before = wc[d];
        __builtin_amdgcn_wave_barrier();
        if (ok && below == cnt - 1u) wc[d] = before + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[it] = before + below;
    }
    __syncthreads();
    // bucket `tid`: wave offsets, the tile's count and start
    uint32_t total = 0;
    if (tid < D) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = s_wcnt[w * D + tid];
            s_wcnt[w * D + tid] = total;
            total += c;
        }
    }
    const uint32_t lstart = block_exclusive_scan<kEncodeThreads>(total, s_shw);
    if (tid < D) s_lstart[tid] = lstart;
    if (tid < nb) {
        tcount[(int64_t)tid * tiles + tile] = total;
        tstart[(int64_t)tid * tiles + tile] = lstart;
        if (total) atomicAdd(&hist[tid], total);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t d = dg[it];
        if (d < (uint32_t)D) {
            const uint32_t lp = s_lstart[d] + s_wcnt[wid * D + d] + rank[it];
            s_keys[lp] = key[it];
            s_vals[lp] = (uint32_t)(base + wid * (64 * kItems) + it * 64 + lane);
        }
    }
    __syncthreads();
    const int64_t rem = n - base;
    const int tile_n = rem < kTile ? (int)rem : kTile;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int q = it * kEncodeThreads + tid;
        if (q < tile_n) {
            kout[base + q] = s_keys[q];
            vout[base + q] = s_vals[q];
        }
    }
}

__global__ __launch_bounds__(256) void clear_u32_kernel(uint32_t *__restrict__ p, int64_t words) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * 256)
        p[i] = 0u;
}

}  // namespace

void launch_bbox(const float2 *pos, const int64_t *sim_off, int num_sims, int max_chunks,
                 float4 *partial, float4 *bbox, uint32_t *slots, hipStream_t s) {
    dim3 grid(max_chunks, num_sims);
    bbox_partial_kernel<<<grid, 256, 0, s>>>(pos, sim_off, max_chunks, partial);
    bbox_final_kernel<<<num_sims, 64, 0, s>>>(sim_off, max_chunks, partial, bbox,
                                              num_sims == 1 ? slots : nullptr);
}

template <int RB>
void launch_encode_rb(const float2 *pos, int64_t n, int morton_mode, float4 *bbox,
                      const uint32_t *slots, const uint8_t *body_sim, uint32_t *keys32,
                      uint64_t *keys64, uint32_t *idx, uint32_t *hist, int passes,
                      hipStream_t s) {
    const unsigned grid = (unsigned)((n + kEncodeTile - 1) / kEncodeTile);
    const int th = kEncodeThreads;
    if (body_sim)
        encode_kernel<1, true, RB>
            <<<grid, th, 0, s>>>(pos, n, bbox, slots, body_sim, keys32, keys64, idx, hist, passes);
    else if (morton_mode == 0)
        encode_kernel<0, false, RB>
            <<<grid, th, 0, s>>>(pos, n, bbox, slots, body_sim, keys32, keys64, idx, hist, passes);
    else
        encode_kernel<1, false, RB>
            <<<grid, th, 0, s>>>(pos, n, bbox, slots, body_sim, keys32, keys64, idx, hist, passes);
}

void launch_encode(const float2 *pos, int64_t n, int morton_mode, float4 *bbox,
                   const uint32_t *slots, const uint8_t *body_sim, uint32_t *keys32,
                   uint64_t *keys64, uint32_t *idx, uint32_t *hist, int passes, int radix_bits,
                   hipStream_t s) {
    if (n <= 0) return;
    if (radix_bits == 11)
        launch_encode_rb<11>(pos, n, morton_mode, bbox, slots, body_sim, keys32, keys64, idx,
                             hist, passes, s);
    else if (radix_bits == 10)
        launch_encode_rb<10>(pos, n, morton_mode, bbox, slots, body_sim, keys32, keys64, idx,
                             hist, passes, s);
    else
        launch_encode_rb<8>(pos, n, morton_mode, bbox, slots, body_sim, keys32, keys64, idx,
                            hist, passes, s);
}

template <int RB, int kItems>
void launch_encode_group_rb(const float2 *pos, int64_t n, int morton_mode, float4 *bbox,
                            const uint32_t *slots, uint32_t *keys32, uint32_t *idx,
                            const uint32_t *split, int nb, uint32_t *hist, uint32_t *tcount,
                            uint32_t *tstart, hipStream_t s) {
    const int tiles = encode_group_tiles(n);
    if (morton_mode == 0)
        encode_group_kernel<0, RB, kItems><<<(unsigned)tiles, kEncodeThreads, 0, s>>>(
            pos, n, bbox, slots, keys32, idx, split, nb, hist, tcount, tstart, tiles);
    else
        encode_group_kernel<1, RB, kItems><<<(unsigned)tiles, kEncodeThreads, 0, s>>>(
            pos, n, bbox, slots, keys32, idx, split, nb, hist, tcount, tstart, tiles);
}

// Tiles of 1024 bodies below 2^18 (a small sort's few tiles would leave most
// CUs idle: 65 536 bodies, encode 12.3 -> 9.1 us; 131 072, 12.8 -> 10.5; no
// gain at 2^18), 4096 above.  Below 2^18 keys there are at most 256 buckets
// (bucket_spacing), which the 1024-body tiles' 8-bit bucket ranking needs.
int encode_group_tile(int64_t n) { return n < (int64_t(1) << 18) ? 1024 : 4096; }
int encode_group_tiles(int64_t n) {
    const int t = encode_group_tile(n);
    return (int)((n + t - 1) / t);
}

void launch_encode_group(const float2 *pos, int64_t n, int morton_mode, float4 *bbox,
                         const uint32_t *slots, uint32_t *keys32, uint32_t *idx,
                         const uint32_t *split, int nb, uint32_t *hist, uint32_t *tcount,
                         uint32_t *tstart, hipStream_t s) {
    if (n <= 0) return;
    // the kernel's tile (kItems x 1024) is encode_group_tile(n); 1024-body
    // tiles come with at most 256 buckets (n < 2^18)
    if (encode_group_tile(n) == 1024)
        launch_encode_group_rb<8, 1>(pos, n, morton_mode, bbox, slots, keys32, idx, split, nb,
                                     hist, tcount, tstart, s);
    else if (nb <= 256)
        launch_encode_group_rb<8, 4>(pos, n, morton_mode, bbox, slots, keys32, idx, split, nb,
                                     hist, tcount, tstart, s);
    else
        launch_encode_group_rb<10, 4>(pos, n, morton_mode, bbox, slots, keys32, idx, split, nb,
                                      hist, tcount, tstart, s);
}

void launch_clear_u32(uint32_t *p, int64_t words, hipStream_t s) {
    if (words <= 0) return;
    int64_t blocks = (words + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    clear_u32_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, words);
}

namespace {
__global__ void fill_f32_kernel(float *p, float value, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}
}  // namespace

void launch_fill_f32(float *p, float value, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    fill_f32_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, value, count);
}

}  // namespace nbs
