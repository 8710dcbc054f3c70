// k_sort.hip -- stable one-sweep LSD radix sort of (key, body index) pairs.
//
// Replaces the reference's webgpu-radix-sort@1.0.7 RadixSortKernel
// (src/simulation.ts:132-140, 400), which sorts the Morton codes with the
// body indices as payload.  The indices start as the identity every step
// (lbvh.wgsl:68), so a stable sort orders equal keys by body index.
//
// Design (Adinets & Merrill, "Onesweep", 2022), written for wave64, with the
// digit width RB = 8, 10 or 11 bits a launch parameter (30-bit keys: 4, 3 or
// 3 passes):
//   0. the encode kernel counts every pass's digits as it writes the keys
//      (k_keys.hip), so the sort starts from complete digit histograms;
//   1. per pass, one kernel: each workgroup scans the pass's 2^RB digit
//      counts into exclusive digit offsets (cheaper than a launch), takes the
//      next tile id from an atomic counter (so it only waits on tiles that
//      already started), ranks its 4096 keys stably with 64-lane ballot
//      matching (RB ballots per key), publishes its per-digit counts,
//      resolves its global offsets by decoupled look-back, and scatters
//      through an LDS staging tile so that stores of one digit run are
//      contiguous;
//   2. the last pass also writes each body's output position (rank), from
//      which the build places the tree's sorted leaf records.
// The histograms, tile counters and look-back words must be zero when a
// sort starts; the tree build that follows every sort clears them for the
// next step (sort_clear_words), so no clearing kernel runs per step.
// Per pass the HBM traffic is one read and one write of keys and values.
#include "nbody_internal.h"

namespace nbs {

namespace {

// Tile size by problem size (sort_tile_for): a pass lasts about one tile's
// load, ranking, look-back and store, so small sorts take small tiles (more
// workgroups at once) and large ones large tiles (shorter look-back
// chains): 2048 keys below 2^16, 4096 up to 2^21, 8192 above (32-bit keys;
// 64-bit keys stop at 4096, as keys, values and the per-wavefront digit
// counts must fit in LDS).  Workgroup size per tile and digit width: 8- and
// 10-bit digits rank 4 or 8 keys per thread (16 wavefronts hide the load
// and ranking latencies), 11-bit digits 8 or 16, in fewer threads, as the
// per-wavefront digit counts (kWaves x 2^RB words) must fit in LDS.
template <int RB, int kTile>
constexpr int sort_threads() {
    return kTile == 2048 ? (RB <= 10 ? 512 : 256)
                         : (kTile == 4096 ? (RB <= 10 ? 1024 : 256) : (RB <= 10 ? 1024 : 512));
}
constexpr int kTileMin = 2048;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr int kLook = 8;  // predecessors read per look-back round trip


// One pass over digit (k >> shift) & (2^RB - 1).  Owner thread t (t <
// kOwners) owns digits [t * DPT, t * DPT + DPT) for the scans and the
// look-back.
//
// kLeaf (the last pass): each item's body also gets its output position,
// rank[body] (a scattered write, which no barrier waits for).
template <typename K, int RB, int kTile, int kThreads, bool kLeaf>
__global__ __launch_bounds__(kThreads) void onesweep_kernel(
    const K *__restrict__ kin, const uint32_t *__restrict__ vin, K *__restrict__ kout,
    uint32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ digit_count,
    uint32_t *lookback, uint32_t *counter, LeafGather lg) {
    constexpr int D = 1 << RB;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kItems = kTile / kThreads;
    constexpr int kOwners = D < kThreads ? D : kThreads;
    constexpr int DPT = D / kOwners;
    constexpr uint32_t kMask = (uint32_t)D - 1u;
    static_assert(kItems * kThreads == kTile, "tile size");
    static_assert(DPT >= 1 && DPT <= 8, "digits per thread");
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wcnt[kWaves * D];  // wave digit counts, then wave offsets
    __shared__ uint32_t s_lstart[D];         // digit start inside the tile
    __shared__ uint32_t s_gstart[D];         // digit start in the output
    __shared__ uint32_t s_shw[kWaves];
    __shared__ K s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    for (int i = tid; i < kWaves * D; i += kThreads) s_wcnt[i] = 0;
    const bool owner = tid < kOwners;
    __syncthreads();  // publishes s_tile
    const uint32_t tile = s_tile;
    const int64_t base = (int64_t)tile * kTile;

    // Wave-striped load: item it of lane l in wave w is tile position
    // w*64*kItems + it*64 + l, so (w, it, l) order is input order.  The
    // pass's digit counts are loaded with the keys: the barriers of their
    // scan wait for both at once.
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
    }
    // where each owned digit starts in the output (exclusive digit scan)
    uint32_t dbase[DPT];
    {
        uint32_t c[DPT], sum = 0;
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            c[j] = owner ? digit_count[tid * DPT + j] : 0u;
            sum += c[j];
        }
        uint32_t run = block_exclusive_scan<kThreads>(sum, s_shw);
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            dbase[j] = run;
            run += c[j];
        }
    }
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const int64_t p = base + (int64_t)wid * (64 * kItems) + it * 64 + lane;
        dg[it] = p < n ? (uint32_t)((k[it] >> shift) & kMask) : (uint32_t)D;
    }

    // Stable rank inside the wave: peers = lanes holding the same digit.
    uint32_t *wc = s_wcnt + wid * D;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t d = dg[it];
        const bool ok = d < (uint32_t)D;
        const uint64_t valid = __ballot(ok);
        uint32_t plo = (uint32_t)valid, phi = (uint32_t)(valid >> 32);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            // m = all ones where the lane's bit b is set; a lane keeps the
            // lanes whose bit agrees with its own: ~(ballot ^ m).
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

    // Owned digits: wave offsets, tile offsets, look-back.
    uint32_t total[DPT], tsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        uint32_t t = 0;
        if (owner) {
            const int d = tid * DPT + j;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = s_wcnt[w * D + d];
                s_wcnt[w * D + d] = t;
                t += c;
            }
        }
        total[j] = t;
        tsum += t;
    }
    {
        uint32_t run = block_exclusive_scan<kThreads>(tsum, s_shw);
        if (owner) {
#pragma unroll
            for (int j = 0; j < DPT; ++j) {
                s_lstart[tid * DPT + j] = run;
                run += total[j];
            }
        }
    }

    uint32_t *mine = lookback + (int64_t)tile * D + tid * DPT;
    uint32_t prefix[DPT];
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        prefix[j] = 0;
        if (owner)
            __hip_atomic_store(&mine[j], (tile == 0 ? kFlagPrefix : kFlagAgg) | total[j],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (tile != 0 && owner) {
        // All owned digits look back together, each from its own
        // predecessor, reading a window of kLook predecessors per round trip
        // (their aggregates are summed in order up to the first inclusive
        // prefix; an unpublished word ends the round and is polled again).
        int64_t pj[DPT];
#pragma unroll
        for (int j = 0; j < DPT; ++j) pj[j] = (int64_t)tile - 1;
        uint32_t pend = (1u << DPT) - 1u;
        for (uint32_t spins = 0; pend; ++spins) {
            if (spins == kSpinLimit) {
                if (lg.fault) atomicOr(lg.fault, kFaultWait);
                break;
            }
            uint32_t w[DPT][kLook];
#pragma unroll
            for (int j = 0; j < DPT; ++j)
#pragma unroll
                for (int q = 0; q < kLook; ++q)
                    w[j][q] = ((pend >> j) & 1u) && pj[j] - q >= 0
                                  ? __hip_atomic_load(&lookback[(pj[j] - q) * D + tid * DPT + j],
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : 0u;
            uint32_t waiting = 0;
#pragma unroll
            for (int j = 0; j < DPT; ++j) {
                if (!((pend >> j) & 1u)) continue;
                int used = 0;
#pragma unroll
                for (int q = 0; q < kLook; ++q) {
                    const uint32_t f = w[j][q] & ~kValMask;
                    if (f == 0u) {
                        waiting = 1;
                        break;
                    }
                    prefix[j] += w[j][q] & kValMask;
                    ++used;
                    if (f == kFlagPrefix) {
                        pend &= ~(1u << j);
                        break;
                    }
                }
                pj[j] -= used;
            }
            if (waiting) __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int j = 0; j < DPT; ++j)
            __hip_atomic_store(&mine[j], kFlagPrefix | (prefix[j] + total[j]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    if (owner) {
#pragma unroll
        for (int j = 0; j < DPT; ++j) s_gstart[tid * DPT + j] = dbase[j] + prefix[j];
    }
    __syncthreads();

    // Scatter into the LDS tile in digit order, then stream out.
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t dd = dg[it];
        if (dd < (uint32_t)D) {
            const uint32_t lp = s_lstart[dd] + s_wcnt[wid * D + dd] + rank[it];
            s_keys[lp] = k[it];
            s_vals[lp] = v[it];
            if constexpr (kLeaf) lg.rank[v[it]] = s_gstart[dd] + (lp - s_lstart[dd]);
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
            const uint32_t dd = (uint32_t)((kk >> shift) & kMask);
            const uint32_t dst = s_gstart[dd] + (uint32_t)(p - (int)s_lstart[dd]);
            kout[dst] = kk;
            vout[dst] = s_vals[p];
        }
    }
}

template <typename K, int RB, int kTile>
int sort_rb_tile(K *keys_a, uint32_t *vals_a, K *keys_b, uint32_t *vals_b, int64_t n, int passes,
                 const SortScratch &sc, const LeafGather &lg, hipStream_t s) {
    constexpr int D = 1 << RB;
    const int64_t tiles = (n + kTile - 1) / kTile;
    K *ki = keys_a, *ko = keys_b;
    uint32_t *vi = vals_a, *vo = vals_b;
    constexpr int T = sort_threads<RB, kTile>();
    for (int p = 0; p < passes; ++p) {
        if (p == passes - 1 && lg.rank)
            onesweep_kernel<K, RB, kTile, T, true><<<(unsigned)tiles, T, 0, s>>>(
                ki, vi, ko, vo, n, RB * p, sc.hist + p * D, sc.lookback + (int64_t)p * tiles * D,
                sc.counters + p, lg);
        else
            onesweep_kernel<K, RB, kTile, T, false><<<(unsigned)tiles, T, 0, s>>>(
                ki, vi, ko, vo, n, RB * p, sc.hist + p * D, sc.lookback + (int64_t)p * tiles * D,
                sc.counters + p, lg);
        K *tk = ki;
        ki = ko;
        ko = tk;
        uint32_t *tv = vi;
        vi = vo;
        vo = tv;
    }
    return passes & 1;
}

template <typename K, int RB>
int sort_rb(K *keys_a, uint32_t *vals_a, K *keys_b, uint32_t *vals_b, int64_t n, int passes,
            const SortScratch &sc, const LeafGather &lg, hipStream_t s) {
    const int tile = sort_tile_for(n, sizeof(K) == 8);
    if (tile == 2048)
        return sort_rb_tile<K, RB, 2048>(keys_a, vals_a, keys_b, vals_b, n, passes, sc, lg, s);
    if constexpr (sizeof(K) == 4)
        if (tile == 8192)
            return sort_rb_tile<K, RB, 8192>(keys_a, vals_a, keys_b, vals_b, n, passes, sc, lg, s);
    return sort_rb_tile<K, RB, 4096>(keys_a, vals_a, keys_b, vals_b, n, passes, sc, lg, s);
}

template <typename K>
int sort_impl(K *keys_a, uint32_t *vals_a, K *keys_b, uint32_t *vals_b, int64_t n, int key_bits,
              int radix_bits, const SortScratch &sc, const LeafGather &lg, hipStream_t s) {
    if (n <= 1) return 0;
    const int passes = sort_passes(key_bits, radix_bits);
    if (radix_bits == 11)
        return sort_rb<K, 11>(keys_a, vals_a, keys_b, vals_b, n, passes, sc, lg, s);
    if (radix_bits == 10)
        return sort_rb<K, 10>(keys_a, vals_a, keys_b, vals_b, n, passes, sc, lg, s);
    return sort_rb<K, 8>(keys_a, vals_a, keys_b, vals_b, n, passes, sc, lg, s);
}

// ---------------------------------------------------------------- bucket sort
// Bodies move little between steps, so the previous step's sorted keys split
// the next step's keys into nearly equal buckets: splitter j (j = 1 .. B-1)
// is the key at sorted position j * S (S about 1024, 2048 or 4096: bucket_spacing).
// A body's bucket is the number of splitters <= its key (the encode finds
// it, so equal keys share a bucket, as does every key of a bucket's range).
// The encode groups each tile's bodies by bucket (encode_group_kernel);
// bucket_sort_kernel gathers each bucket tile after tile, one workgroup per
// bucket, and sorts it with stable 8-bit LSD passes over the key bits that
// vary inside the bucket's key range, or by counting within bins.  A sort of
// at most kSortLdsMax keys is one bucket, sorted by one workgroup straight
// from the encode.  The result is the radix sort's: the stable order is
// unique.
constexpr int kBucketThreads = 1024;
constexpr int kBucketWaves = kBucketThreads / kWave;
constexpr int kBucketRB = 8;
constexpr int kBucketD = 1 << kBucketRB;
constexpr int kBinBits = 11;  // bins of the in-LDS bucket sort
constexpr int kBins = 1 << kBinBits;
constexpr int kBinMax = 32;   // largest bin ranked by counting
// Counting beats the LSD passes when those are four or more (wide key
// ranges: the few buckets of small sorts); at three (the 4096-key buckets
// of a million bodies) the passes are a little faster (22.6 against
// 24.5 us at 800k).
constexpr int kBinMinPasses = 4;

// Stable rank of each of the wavefront's first `rows` items among its items
// of the same digit (dg == kBucketD: an empty slot), counting into the
// wavefront's digit counters wc (kBucketD words, zero before the first
// item): ballot matching as in the one-sweep passes.  `rows` is uniform.
template <int kItems>
__device__ __forceinline__ void wave_rank8(const uint32_t (&dg)[kItems], uint32_t (&rank)[kItems],
                                           uint32_t *wc, int rows) {
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        if (it >= rows) break;
        const uint32_t d = dg[it];
        const bool ok = d < (uint32_t)kBucketD;
        const uint64_t valid = __ballot(ok);
        uint32_t plo = (uint32_t)valid, phi = (uint32_t)(valid >> 32);
#pragma unroll
        for (int b = 0; b < kBucketRB; ++b) {
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
        if (ok && (below == cnt - 1u)) wc[d] = before + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[it] = before + below;
    }
}

// After wave_rank8 in every wavefront (and a barrier): s_wcnt[w][d] becomes
// the count of digit d in the wavefronts before w, s_start[d] the first slot
// of digit d, s_total[d] its count.  Ends with a barrier.
__device__ __forceinline__ void digit_starts(uint32_t *s_wcnt, uint32_t *s_start,
                                             uint32_t *s_total, uint32_t *s_shw) {
    const int tid = threadIdx.x;
    uint32_t t = 0;
    if (tid < kBucketD) {
        uint32_t c[kBucketWaves];
#pragma unroll
        for (int w = 0; w < kBucketWaves; ++w) c[w] = s_wcnt[w * kBucketD + tid];
#pragma unroll
        for (int w = 0; w < kBucketWaves; ++w) {
            s_wcnt[w * kBucketD + tid] = t;
            t += c[w];
        }
    }
    const uint32_t start = block_exclusive_scan<kBucketThreads>(t, s_shw);
    if (tid < kBucketD) {
        s_start[tid] = start;
        s_total[tid] = t;
    }
    __syncthreads();
}

// One workgroup per bucket (blockIdx.x), kCap keys at most in LDS (larger
// buckets go through memory in chunks of kCap).  Bucket b starts at
// position start[b] and holds count[b] keys in [split[b-1], split[b] - 1]
// (split[-1] = 0, split[B-1] = the largest key); start == null: one bucket
// of n keys.  Input (kin, vin) and output (kout, vout) cover the same
// positions; a bucket too large for LDS also overwrites its input.
//
// kGather: the input is grouped by bucket inside tiles of bg.tile positions
// (encode_group_kernel); bucket b's keys are gathered tile after tile, at
// tile t's positions [t * bg.tile + tstart[b * tiles + t], + tcount[b * tiles
// + t]), which is input order; the bucket starts at the sum of the counts
// of the buckets before it.  A bucket too large for LDS is gathered into
// its output positions and sorted there, through (sk2, sv2) at the same
// positions.
struct BucketGather {
    const uint32_t *tcount = nullptr, *tstart = nullptr;
    int tiles = 0, tile = 0;  // tiles of `tile` positions
    uint32_t *sk2 = nullptr, *sv2 = nullptr;
};
constexpr int kGMaxTiles = 1024;
template <int kCap, bool kGather = false>
__global__ __launch_bounds__(kBucketThreads) void bucket_sort_kernel(
    uint32_t *kin, uint32_t *vin, uint32_t *kout, uint32_t *vout, int64_t n, uint32_t key_max,
    int nb, const uint32_t *__restrict__ start, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ split, uint32_t *__restrict__ rank_out, BucketGather bg = {}) {
    constexpr int kItems = kCap / kBucketThreads;
    static_assert(kItems * kBucketThreads == kCap, "bucket capacity");
    __shared__ uint32_t s_k[kCap];
    __shared__ uint32_t s_v[kCap];
    __shared__ uint32_t s_wcnt[kBucketWaves * kBucketD];
    __shared__ uint32_t s_start[kBucketD], s_total[kBucketD], s_run[kBucketD];
    __shared__ uint32_t s_shw[kBucketWaves];
    __shared__ uint32_t s_seg[kGather ? kGMaxTiles : 1], s_tst[kGather ? kGMaxTiles : 1];
    __shared__ uint32_t s_off;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int b = (int)blockIdx.x;

    uint32_t off, cnt;
    if constexpr (kGather) {
        // the bucket's start (counts of the buckets before it) and, per tile,
        // where its keys begin among the bucket's
        cnt = count[b];
        if (cnt == 0u) return;
        uint32_t before = 0;
        for (int i = tid; i < b; i += kBucketThreads) before += count[i];
        uint32_t h = 0;
        if (tid < bg.tiles) {
            h = bg.tcount[(int64_t)b * bg.tiles + tid];
            s_tst[tid] = bg.tstart[(int64_t)b * bg.tiles + tid];
        }
        const uint32_t seg = block_exclusive_scan<kBucketThreads>(h, s_shw);
        if (tid < bg.tiles) s_seg[tid] = seg;
        off = block_exclusive_scan<kBucketThreads>(before, s_shw);
        if (tid == kBucketThreads - 1) s_off = off + before;  // the total
        __syncthreads();
        off = s_off;
    } else {
        off = start ? start[b] : 0u;
        cnt = start ? count[b] : (uint32_t)n;
        if (cnt == 0u) return;
    }
    // position p (< cnt) of the bucket's gathered input: the last tile whose
    // segment starts at or before p
    auto src = [&](uint32_t p) {
        int t = 0;
        for (int step = kGMaxTiles / 2; step >= 1; step >>= 1)
            if (t + step < bg.tiles && s_seg[t + step] <= p) t += step;
        return (uint32_t)t * (uint32_t)bg.tile + s_tst[t] + (p - s_seg[t]);
    };
    // bits that vary inside the bucket's key range
    const uint32_t klo = b == 0 ? 0u : split[b - 1];
    const uint32_t khi = b == nb - 1 ? key_max : split[b] - 1u;
    const uint32_t diff = klo ^ khi;
    const int passes = diff ? (32 - __clz((int)diff) + kBucketRB - 1) / kBucketRB : 0;

    uint32_t k[kItems], v[kItems], dg[kItems], rk[kItems];
    if (cnt <= (uint32_t)kCap) {
        // item it < rows of lane l in wave w: bucket position
        // w*64*rows + it*64 + l (every wavefront gets the same rows)
        const int rows = (int)((cnt + kBucketThreads - 1) / kBucketThreads);
        auto slot = [&](int it) { return (uint32_t)(wid * 64 * rows + it * 64 + lane); };
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            if (it >= rows) break;
            const uint32_t p = slot(it);
            if constexpr (kGather) {
                const uint32_t q = p < cnt ? src(p) : 0u;
                k[it] = p < cnt ? kin[q] : 0u;
                v[it] = p < cnt ? vin[q] : 0u;
            } else {
                k[it] = p < cnt ? kin[off + p] : 0u;
                v[it] = p < cnt ? vin[off + p] : 0u;
            }
        }
        // Bins: the top (up to kBinBits) of the bits that vary in the
        // bucket, so a bin is a key interval and bins are in key order.
        // Each key goes to its bin (any order), then takes its place in
        // the bin by counting the bin's keys ordered before it (key, then
        // body index: the stable order).  Bins of more than kBinMax keys
        // (clustered or repeated keys) take the LSD passes instead.
        const int vbits = diff ? 32 - __clz((int)diff) : 0;
        const int bbits = vbits < kBinBits ? vbits : kBinBits;
        const int bsh = vbits - bbits;
        const uint32_t bmask = (1u << bbits) - 1u;
        uint32_t *s_bcnt = s_wcnt, *s_bstart = s_wcnt + kBins;
        for (int i = tid; i < kBins; i += kBucketThreads) s_bcnt[i] = 0u;
        if (tid == 0) s_run[0] = 0u;  // the largest bin
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            if (it >= rows) break;
            if (slot(it) < cnt) atomicAdd(&s_bcnt[(k[it] >> bsh) & bmask], 1u);
        }
        __syncthreads();
        uint32_t big = 0;  // the largest bin's count (when above kBinMax)
        {
            // kBins / kBucketThreads bins per thread: exclusive scan, largest bin
            constexpr int kPer = kBins / kBucketThreads;
            uint32_t c[kPer], sum = 0;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                c[j] = s_bcnt[tid * kPer + j];
                sum += c[j];
                big = c[j] > big ? c[j] : big;
            }
            uint32_t run = block_exclusive_scan<kBucketThreads>(sum, s_shw);
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                s_bstart[tid * kPer + j] = run;
                s_bcnt[tid * kPer + j] = 0u;  // now the bins' fill cursors
                run += c[j];
            }
            if (big > (uint32_t)kBinMax) atomicMax(&s_run[0], big);
            __syncthreads();
            big = s_run[0];
        }
        if (big <= (uint32_t)kBinMax && passes >= kBinMinPasses) {
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                if (it >= rows) break;
                if (slot(it) < cnt) {
                    const uint32_t d = (k[it] >> bsh) & bmask;
                    const uint32_t at = s_bstart[d] + atomicAdd(&s_bcnt[d], 1u);
                    s_k[at] = k[it];
                    s_v[at] = v[it];
                }
            }
            __syncthreads();
            for (uint32_t p = (uint32_t)tid; p < cnt; p += kBucketThreads) {
                const uint32_t kp = s_k[p], vp = s_v[p];
                const uint32_t d = (kp >> bsh) & bmask;
                const uint32_t b0 = s_bstart[d], b1 = b0 + s_bcnt[d];
                uint32_t r = 0;
                for (uint32_t q = b0; q < b1; ++q) {
                    const uint32_t kq = s_k[q];
                    r += (kq < kp || (kq == kp && s_v[q] < vp)) ? 1u : 0u;
                }
                kout[off + b0 + r] = kp;
                vout[off + b0 + r] = vp;
                rank_out[vp] = off + b0 + r;
            }
            return;
        }
        __syncthreads();  // the bins' counters are the passes' counters
        for (int q = 0; q < passes; ++q) {
            const int shift = kBucketRB * q;
            for (int i = tid; i < kBucketWaves * kBucketD; i += kBucketThreads) s_wcnt[i] = 0u;
            __syncthreads();
#pragma unroll
            for (int it = 0; it < kItems; ++it)
                dg[it] = it < rows && slot(it) < cnt ? (k[it] >> shift) & (kBucketD - 1u)
                                                     : (uint32_t)kBucketD;
            wave_rank8<kItems>(dg, rk, s_wcnt + wid * kBucketD, rows);
            __syncthreads();
            digit_starts(s_wcnt, s_start, s_total, s_shw);
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                if (it >= rows) break;
                if (dg[it] < (uint32_t)kBucketD) {
                    const uint32_t lp = s_start[dg[it]] + s_wcnt[wid * kBucketD + dg[it]] + rk[it];
                    s_k[lp] = k[it];
                    s_v[lp] = v[it];
                    // the last pass knows each body's place: its scattered
                    // write overlaps the rest of the kernel
                    if (q == passes - 1) rank_out[v[it]] = off + lp;
                }
            }
            __syncthreads();
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                if (it >= rows) break;
                const uint32_t p = slot(it);
                if (p < cnt) {
                    k[it] = s_k[p];
                    v[it] = s_v[p];
                }
            }
        }
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            if (it >= rows) break;
            const uint32_t p = slot(it);
            if (p < cnt) {
                kout[off + p] = k[it];
                vout[off + p] = v[it];
                if (passes == 0) rank_out[v[it]] = off + p;
            }
        }
        return;
    }

    // A bucket larger than kCap: each pass counts its digits over the whole
    // bucket, then moves it chunk by chunk in input order (stable) between
    // the input and output positions.
    auto slot = [&](int it) { return (uint32_t)(wid * (64 * kItems) + it * 64 + lane); };
    uint32_t *sk = kin + off, *sv = vin + off, *dk = kout + off, *dv = vout + off;
    if constexpr (kGather) {
        // gathered into the output positions first, then sorted between them
        // and the same positions of (sk2, sv2)
        for (uint32_t p = (uint32_t)tid; p < cnt; p += kBucketThreads) {
            const uint32_t q = src(p);
            kout[off + p] = kin[q];
            vout[off + p] = vin[q];
        }
        __syncthreads();
        sk = kout + off;
        sv = vout + off;
        dk = bg.sk2 + off;
        dv = bg.sv2 + off;
    }
    for (int q = 0; q < passes; ++q) {
        const int shift = kBucketRB * q;
        if (tid < kBucketD) s_run[tid] = 0u;
        __syncthreads();
        for (uint32_t c0 = 0; c0 < cnt; c0 += (uint32_t)kCap)
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                const uint32_t p = c0 + slot(it);
                if (p < cnt) atomicAdd(&s_run[(sk[p] >> shift) & (kBucketD - 1u)], 1u);
            }
        __syncthreads();
        {
            const uint32_t t = tid < kBucketD ? s_run[tid] : 0u;
            const uint32_t first = block_exclusive_scan<kBucketThreads>(t, s_shw);
            if (tid < kBucketD) s_run[tid] = first;
            __syncthreads();
        }
        for (uint32_t c0 = 0; c0 < cnt; c0 += (uint32_t)kCap) {
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                const uint32_t p = c0 + slot(it);
                k[it] = p < cnt ? sk[p] : 0u;
                v[it] = p < cnt ? sv[p] : 0u;
                dg[it] = p < cnt ? (k[it] >> shift) & (kBucketD - 1u) : (uint32_t)kBucketD;
            }
            for (int i = tid; i < kBucketWaves * kBucketD; i += kBucketThreads) s_wcnt[i] = 0u;
            __syncthreads();
            wave_rank8<kItems>(dg, rk, s_wcnt + wid * kBucketD, kItems);
            __syncthreads();
            digit_starts(s_wcnt, s_start, s_total, s_shw);
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                if (dg[it] < (uint32_t)kBucketD) {
                    const uint32_t dst = s_run[dg[it]] + s_wcnt[wid * kBucketD + dg[it]] + rk[it];
                    dk[dst] = k[it];
                    dv[dst] = v[it];
                }
            }
            __syncthreads();  // the chunk's stores are done before the next chunk's loads
            if (tid < kBucketD) s_run[tid] += s_total[tid];
            __syncthreads();
        }
        uint32_t *tk = sk, *tv = sv;
        sk = dk;
        sv = dv;
        dk = tk;
        dv = tv;
    }
    // the sorted bucket is at (sk, sv): move it to the output if it is not
    // there already, and record every body's position
    const bool moved = sk != kout + off;
    for (uint32_t p = (uint32_t)tid; p < cnt; p += kBucketThreads) {
        const uint32_t kk = sk[p], vv = sv[p];
        if (moved) {
            kout[off + p] = kk;
            vout[off + p] = vv;
        }
        rank_out[vv] = off + p;
    }
}

}  // namespace

int sort_tile_for(int64_t n, bool wide_keys) {
    if (n < (int64_t(1) << 16)) return 2048;
    if (wide_keys || n <= (int64_t(1) << 21)) return 4096;
    return 8192;
}

// the most tiles a sort of n keys can have (the smallest tile)
int64_t sort_tiles(int64_t n) { return (n + kTileMin - 1) / kTileMin; }

int sort_passes(int key_bits, int radix_bits) { return (key_bits + radix_bits - 1) / radix_bits; }

// Histograms (kSortMaxPasses x 2^11 words), 64 counter words, then the
// look-back words of the widest configuration: 4 passes of 2^11 digits
// (38-bit ensemble keys at 11 bits) per tile.
int64_t sort_scratch_words(int64_t cap) {
    return kSortHistWords + 64 + (int64_t)4 * 2048 * sort_tiles(cap);
}

int64_t sort_clear_words(int64_t n, int key_bits, int radix_bits, bool wide_keys) {
    const int tile = sort_tile_for(n, wide_keys);
    return kSortHistWords + 64 + (int64_t)sort_passes(key_bits, radix_bits) *
                                     ((n + tile - 1) / tile) * ((int64_t)1 << radix_bits);
}

void sort_bind_scratch(SortScratch &sc, uint32_t *base, int64_t cap) {
    sc.hist = base;
    sc.counters = base + kSortHistWords;
    sc.lookback = sc.counters + 64;
    sc.words = sort_scratch_words(cap);
}

int launch_sort_u32(uint32_t *keys_a, uint32_t *vals_a, uint32_t *keys_b, uint32_t *vals_b,
                    int64_t n, int key_bits, int radix_bits, const SortScratch &sc,
                    const LeafGather &lg, hipStream_t s) {
    return sort_impl<uint32_t>(keys_a, vals_a, keys_b, vals_b, n, key_bits, radix_bits, sc, lg, s);
}

uint32_t sort_key_max(int key_bits) {
    return key_bits >= 32 ? 0xFFFFFFFFu : (1u << key_bits) - 1u;
}

int launch_sort_small(uint32_t *keys_a, uint32_t *vals_a, uint32_t *keys_b, uint32_t *vals_b,
                      int64_t n, int key_bits, const LeafGather &lg, hipStream_t s) {
    bucket_sort_kernel<kBucketCap><<<1, kBucketThreads, 0, s>>>(
        keys_a, vals_a, keys_b, vals_b, n, sort_key_max(key_bits), 1, nullptr, nullptr, nullptr,
        lg.rank);
    return 1;
}

int launch_sort_gathered(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out,
                         uint32_t *vals_out, uint32_t *keys_spill, uint32_t *vals_spill, int64_t n,
                         int key_bits, int nb, const uint32_t *split, const uint32_t *hist,
                         const uint32_t *tcount, const uint32_t *tstart, const LeafGather &lg,
                         hipStream_t s) {
    BucketGather bg;
    bg.tcount = tcount;
    bg.tstart = tstart;
    bg.tile = encode_group_tile(n);
    bg.tiles = encode_group_tiles(n);
    bg.sk2 = keys_spill;
    bg.sv2 = vals_spill;
    if (bg.tiles > kGMaxTiles) return -1;
    bucket_sort_kernel<kBucketCap, true><<<(unsigned)nb, kBucketThreads, 0, s>>>(
        const_cast<uint32_t *>(keys_in), const_cast<uint32_t *>(vals_in), keys_out, vals_out, n,
        sort_key_max(key_bits), nb, nullptr, hist, split, lg.rank, bg);
    return 0;
}

int launch_sort_u64(uint64_t *keys_a, uint32_t *vals_a, uint64_t *keys_b, uint32_t *vals_b,
                    int64_t n, int key_bits, int radix_bits, const SortScratch &sc,
                    const LeafGather &lg, hipStream_t s) {
    return sort_impl<uint64_t>(keys_a, vals_a, keys_b, vals_b, n, key_bits, radix_bits, sc, lg, s);
}

}  // namespace nbs
