// k_sort.hip -- stable one-sweep LSD radix sort of (key, body index) pairs.
//
// Replaces the reference's webgpu-radix-sort@1.0.7 RadixSortKernel
// (src/simulation.ts:132-140, 400), which sorts the Morton codes with the
// body indices as payload.  The indices start as the identity every step
// (lbvh.wgsl:68), so a stable sort orders equal keys by body index.
//
// Design (Adinets & Merrill, "Onesweep", 2022), written for wave64:
//   1. one histogram kernel counts every pass's 8-bit digits at once
//      (LDS histograms, one global atomic per bin per workgroup);
//   2. per pass, one kernel: each workgroup scans the pass's digit counts
//      into exclusive digit offsets (256 entries, cheaper than a launch),
//      takes the next tile id from an atomic counter (so it only waits on
//      tiles that already started),
//      ranks its 4096 keys stably with 64-lane ballot matching, publishes
//      its per-digit counts, resolves its global offsets by decoupled
//      look-back, and scatters through an LDS staging tile so that stores
//      of one digit run are contiguous.
// Per pass the HBM traffic is one read and one write of keys and values.
#include "nbody_internal.h"

namespace nbs {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
static_assert(kThreads == 256, "one thread per 8-bit digit");
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr int kMaxPasses = 8;

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh4) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh4[wid] = x;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < kWaves; ++w)
        if (w < wid) before += sh4[w];
    __syncthreads();
    return before + x - v;
}

template <typename K>
__global__ __launch_bounds__(kThreads) void hist_kernel(const K *__restrict__ keys, int64_t n,
                                                        int passes, uint32_t *__restrict__ hist) {
    __shared__ uint32_t sh[kMaxPasses * 256];
    for (int i = threadIdx.x; i < passes * 256; i += kThreads) sh[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const K k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&sh[p * 256 + (uint32_t)((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += kThreads)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

template <typename K>
__global__ __launch_bounds__(kThreads) void onesweep_kernel(
    const K *__restrict__ kin, const uint32_t *__restrict__ vin, K *__restrict__ kout,
    uint32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ digit_count,
    uint32_t *lookback, uint32_t *counter) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wcnt[kWaves * 256];  // wave digit counts, then wave offsets
    __shared__ uint32_t s_lstart[256];         // digit start inside the tile
    __shared__ uint32_t s_gstart[256];         // digit start in the output
    __shared__ uint32_t s_sh4[kWaves];
    __shared__ K s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    for (int i = tid; i < kWaves * 256; i += kThreads) s_wcnt[i] = 0;
    // thread d: where digit d starts in the output (exclusive digit scan)
    const uint32_t digit_base = block_exclusive_scan(digit_count[tid], s_sh4);
    const uint32_t tile = s_tile;
    const int64_t base = (int64_t)tile * kTile;

    // Wave-striped load: item it of lane l in wave w is tile position
    // w*64*kItems + it*64 + l, so (w, it, l) order is input order.
    K k[kItems];
    uint32_t v[kItems];
    uint32_t dg[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t p = base + (int64_t)wid * (64 * kItems) + it * 64 + lane;
        const bool ok = p < n;
        k[it] = ok ? kin[p] : (K)0;
        v[it] = ok ? vin[p] : 0u;
        dg[it] = ok ? (uint32_t)((k[it] >> shift) & 255u) : 256u;
    }

    // Stable rank inside the wave: peers = lanes holding the same digit.
    uint32_t *wc = s_wcnt + wid * 256;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t d = dg[it];
        const bool ok = d < 256u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
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

    // Thread d owns digit d: wave offsets, tile offsets, look-back.
    const uint32_t d = tid;
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wcnt[w * 256 + d];
        s_wcnt[w * 256 + d] = total;
        total += c;
    }
    const uint32_t lstart = block_exclusive_scan(total, s_sh4);
    s_lstart[d] = lstart;

    uint32_t *mine = lookback + (int64_t)tile * 256;
    uint32_t prefix = 0;
    if (tile == 0) {
        __hip_atomic_store(&mine[d], kFlagPrefix | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(&mine[d], kFlagAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t j = (int64_t)tile - 1;
        for (;;) {
            const uint32_t w = __hip_atomic_load(&lookback[j * 256 + d], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t f = w & ~kValMask;
            if (f == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += w & kValMask;
            if (f == kFlagPrefix) break;
            --j;
        }
        __hip_atomic_store(&mine[d], kFlagPrefix | (prefix + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gstart[d] = digit_base + prefix;
    __syncthreads();

    // Scatter into the LDS tile in digit order, then stream out.
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t dd = dg[it];
        if (dd < 256u) {
            const uint32_t lp = s_lstart[dd] + s_wcnt[wid * 256 + dd] + rank[it];
            s_keys[lp] = k[it];
            s_vals[lp] = v[it];
        }
    }
    __syncthreads();
    const int64_t rem = n - base;
    const int tile_n = rem < kTile ? (int)rem : kTile;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int p = it * kThreads + tid;
        if (p < tile_n) {
            const K kk = s_keys[p];
            const uint32_t dd = (uint32_t)((kk >> shift) & 255u);
            const uint32_t dst = s_gstart[dd] + (uint32_t)(p - (int)s_lstart[dd]);
            kout[dst] = kk;
            vout[dst] = s_vals[p];
        }
    }
}

template <typename K>
int sort_impl(K *keys_a, uint32_t *vals_a, K *keys_b, uint32_t *vals_b, int64_t n, int bits,
              const SortScratch &sc, hipStream_t s) {
    if (n <= 1) return 0;
    const int passes = (bits + 7) / 8;
    const int64_t tiles = sort_tiles(n);
    launch_clear_u32(sc.hist, sc.words, s);
    int64_t hblocks = tiles < 1024 ? tiles : 1024;
    hist_kernel<K><<<(unsigned)hblocks, kThreads, 0, s>>>(keys_a, n, passes, sc.hist);
    K *ki = keys_a, *ko = keys_b;
    uint32_t *vi = vals_a, *vo = vals_b;
    for (int p = 0; p < passes; ++p) {
        onesweep_kernel<K><<<(unsigned)tiles, kThreads, 0, s>>>(
            ki, vi, ko, vo, n, 8 * p, sc.hist + p * 256, sc.lookback + (int64_t)p * tiles * 256,
            sc.counters + p);
        K *tk = ki;
        ki = ko;
        ko = tk;
        uint32_t *tv = vi;
        vi = vo;
        vo = tv;
    }
    return passes & 1;
}

}  // namespace

int64_t sort_tiles(int64_t n) { return (n + kTile - 1) / kTile; }

int64_t sort_scratch_words(int64_t n, int passes) {
    // hist and counters live in front of the look-back words; the clear
    // kernel zeroes all three in one launch.
    return (int64_t)passes * 256 + (int64_t)passes * sort_tiles(n) * 256 + 64;
}

void sort_bind_scratch(SortScratch &sc, uint32_t *base, int64_t n, int passes) {
    sc.hist = base;
    sc.counters = base + (int64_t)passes * 256;
    sc.lookback = sc.counters + 64;
    sc.words = sort_scratch_words(n, passes);
}

int launch_sort_u32(uint32_t *keys_a, uint32_t *vals_a, uint32_t *keys_b, uint32_t *vals_b,
                    int64_t n, int bits, const SortScratch &sc, hipStream_t s) {
    return sort_impl<uint32_t>(keys_a, vals_a, keys_b, vals_b, n, bits, sc, s);
}

int launch_sort_u64(uint64_t *keys_a, uint32_t *vals_a, uint64_t *keys_b, uint32_t *vals_b,
                    int64_t n, int bits, const SortScratch &sc, hipStream_t s) {
    return sort_impl<uint64_t>(keys_a, vals_a, keys_b, vals_b, n, bits, sc, s);
}

}  // namespace nbs
