// k_direct.hip -- FP64 softened direct sum, the accuracy reference of the walk.
//
// The reference approximates every body's acceleration with the Barnes-Hut
// walk (src/shaders/compute/barnes_hut.wgsl:48-111); its error is measured
// against the exact O(N^2) sum of the same softened pull,
//   a_i = sum_{j != i} G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^{3/2},
// in binary64, with eps constant or the reference's per-pair
// epsMult * sqrt(G * max(m_i, m_j) * dt) (barnes_hut.wgsl:90-91).
//
// Layout: a workgroup owns 256 targets (one per thread, in registers) and one
// contiguous chunk of the sources, which it streams through LDS 256 at a
// time; each (target block, source chunk) pair writes its partial sums, and a
// second kernel adds the chunks in chunk order.  Every source is read once per
// target block, so the sum is bound by FP64 issue (the square root and the
// division), not by memory.
#include "nbody_internal.h"

namespace nbs {

namespace {

constexpr int kDirectBlock = 256;

__global__ __launch_bounds__(kDirectBlock) void direct_partial_kernel(
    const float2 *__restrict__ pos, const float *__restrict__ mass, int64_t n,
    const int64_t *__restrict__ targets, int64_t ntargets, int64_t chunk, StepParams p,
    double2 *__restrict__ partial) {
    __shared__ float2 sp[kDirectBlock];
    __shared__ float sm[kDirectBlock];
    const int64_t k = (int64_t)blockIdx.x * kDirectBlock + threadIdx.x;  // target slot
    const bool active = k < ntargets;
    const int64_t i = active ? targets[k] : -1;
    double xi = 0.0, yi = 0.0, mi = 0.0;
    if (active) {
        xi = pos[i].x;
        yi = pos[i].y;
        mi = mass[i];
    }
    const double g = p.grav, dt = p.dt, mult = p.eps_mult;
    const double eps2c = (double)p.eps * (double)p.eps;
    const bool ref_soft = p.soft_mode == 0;
    const int64_t j0 = (int64_t)blockIdx.y * chunk;
    const int64_t j1 = min(n, j0 + chunk);
    double ax = 0.0, ay = 0.0;
    for (int64_t base = j0; base < j1; base += kDirectBlock) {
        const int64_t j = base + threadIdx.x;
        __syncthreads();
        if (j < j1) {
            sp[threadIdx.x] = pos[j];
            sm[threadIdx.x] = mass[j];
        }
        __syncthreads();
        const int cnt = (int)min<int64_t>(kDirectBlock, j1 - base);
        for (int q = 0; q < cnt; ++q) {
            const double rx = (double)sp[q].x - xi, ry = (double)sp[q].y - yi;
            const double mj = sm[q];
            double eps2 = eps2c;
            if (ref_soft) {
                const double e = mult * sqrt(g * (mi > mj ? mi : mj) * dt);
                eps2 = e * e;
            }
            const double inv = 1.0 / sqrt(rx * rx + ry * ry + eps2);
            const double f = (base + q == i) ? 0.0 : g * mj * inv * inv * inv;
            ax += f * rx;
            ay += f * ry;
        }
    }
    if (active) partial[(size_t)blockIdx.y * (size_t)ntargets + (size_t)k] = make_double2(ax, ay);
}

__global__ void direct_reduce_kernel(const double2 *__restrict__ partial, int64_t ntargets,
                                     int chunks, double2 *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntargets) return;
    double ax = 0.0, ay = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const double2 v = partial[(size_t)c * (size_t)ntargets + (size_t)k];
        ax += v.x;
        ay += v.y;
    }
    out[k] = make_double2(ax, ay);
}

}  // namespace

// Source chunks per target block: enough workgroups to fill the GPU.
int direct_chunks(int64_t n, int64_t ntargets) {
    const int64_t blocks = (ntargets + kDirectBlock - 1) / kDirectBlock;
    int64_t chunks = (4096 + blocks - 1) / blocks;
    const int64_t max_chunks = (n + kDirectBlock - 1) / kDirectBlock;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks > 65535) chunks = 65535;
    return (int)(chunks < 1 ? 1 : chunks);
}

void launch_direct_f64(const float2 *pos, const float *mass, int64_t n, const int64_t *targets,
                       int64_t ntargets, const StepParams &p, double2 *partial, int chunks,
                       double2 *out, hipStream_t s) {
    if (ntargets <= 0 || n <= 0) return;
    const int64_t chunk = ((n + chunks - 1) / chunks + kDirectBlock - 1) / kDirectBlock * kDirectBlock;
    dim3 grid((unsigned)((ntargets + kDirectBlock - 1) / kDirectBlock), (unsigned)chunks);
    direct_partial_kernel<<<grid, kDirectBlock, 0, s>>>(pos, mass, n, targets, ntargets, chunk, p,
                                                        partial);
    direct_reduce_kernel<<<(unsigned)((ntargets + 255) / 256), 256, 0, s>>>(partial, ntargets,
                                                                            chunks, out);
}

}  // namespace nbs
