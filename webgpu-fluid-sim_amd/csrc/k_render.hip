// k_render.hip -- density splat and tone map: the reference's Renderer frame
// (renderer.ts:293-329), which draws density.wgsl (one 2R x 2R quad per body,
// R = 2 * radiusMultiplier pixels, additive blending into an r16float
// texture) and then tone_map.wgsl (a full-screen pass: exposure, Reinhard,
// gamma, a four-stop colour ramp), as compute kernels.
//
//   splat_kernel     one thread per body; for every pixel centre of its quad
//                    with r = |uv| <= 1 it adds base * exp(-10 r^2) *
//                    zoom_scale (density.wgsl:77-96) to that pixel as a 64-bit
//                    fixed-point number (units of 2^-40) with an integer
//                    atomic, so the sum does not depend on the order of the
//                    additions (the reference rounds every blend to fp16 in
//                    draw order; here the exact sum is rounded to fp16 once).
//   tone_lut_kernel  the tone map of every fp16 density value, in binary64
//                    with the shader's constants, quantised to 8-bit RGBA:
//                    the per-pixel pass is then one table lookup.
//   tone_map_kernel  per output pixel: the density texel (the full-screen
//                    triangle's uv flips rows: output row y reads texture row
//                    H-1-y, tone_map.wgsl:58-70), fixed point -> fp16 (round
//                    to nearest even) -> table -> RGBA.
// Every floating-point step is a single binary32 operation in the order the
// shader writes it (no contraction), and exp is evaluated by render_exp, so
// the CPU oracle reproduces the images bit for bit.
#include "nbody_internal.h"

namespace nbs {

namespace {

constexpr double kFixedScale = 1099511627776.0;  // 2^40

// exp(x) for x <= 0 in binary32: x log2(e) = n + f with n = rint, |f| <= 1/2,
// 2^f by its degree-7 Taylor polynomial in f ln 2 (Horner), then * 2^n.
// Relative error below 1e-6 on [-10, 0] (the splat's range).
__device__ __forceinline__ float render_exp(float x) {
    const float t = x * 1.44269504f;
    const float n = rintf(t);
    const float f = t - n;
    float p = 1.5252734e-05f;
    p = p * f + 1.5403530e-04f;
    p = p * f + 1.3333558e-03f;
    p = p * f + 9.6181291e-03f;
    p = p * f + 5.5504109e-02f;
    p = p * f + 2.4022651e-01f;
    p = p * f + 6.9314718e-01f;
    p = p * f + 1.0f;
    return ldexpf(p, (int)n);
}

__device__ __forceinline__ uint64_t to_fixed(float v) {
    return (uint64_t)rint((double)v * kFixedScale);
}

__global__ __launch_bounds__(256) void splat_kernel(const float2 *__restrict__ pos,
                                                    const float *__restrict__ radius, uint32_t n,
                                                    RenderParams rp,
                                                    unsigned long long *__restrict__ dens) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 p = pos[i];
    const float rm = radius[i];
    const float sx = 2.0f / (float)rp.width, sy = 2.0f / (float)rp.height;  // density.wgsl:62
    const float cnx = (p.x - rp.cam_x) / rp.half_x;                        // density.wgsl:35-37
    const float cny = (p.y - rp.cam_y) / rp.half_y;
    const float rpx = 2.0f * rm;                                           // density.wgsl:61
    const float rnx = rpx * sx, rny = rpx * sy;                            // density.wgsl:63
    if (!(fabsf(cnx) < 1e30f) || !(fabsf(cny) < 1e30f) || !(rnx > 0.0f) || !(rny > 0.0f) ||
        !(rnx < 1e30f))
        return;
    const float base = rm > 1.0f ? 0.1f : 0.02f;                           // density.wgsl:89-92
    // pixel centres x_ndc = (x + 0.5) sx - 1, y_ndc = 1 - (y + 0.5) sy; the
    // range is conservative, the r <= 1 test below decides
    const float fx0 = floorf((cnx - rnx + 1.0f) / sx - 0.5f) - 1.0f;
    const float fx1 = ceilf((cnx + rnx + 1.0f) / sx - 0.5f) + 1.0f;
    const float fy0 = floorf((1.0f - cny - rny) / sy - 0.5f) - 1.0f;
    const float fy1 = ceilf((1.0f - cny + rny) / sy - 0.5f) + 1.0f;
    // clamped to [-1, size] in float before the conversion to int
    const float W = (float)rp.width, H = (float)rp.height;
    const int x0 = (int)fminf(fmaxf(fx0, 0.0f), W), x1 = (int)fmaxf(fminf(fx1, W - 1.0f), -1.0f);
    const int y0 = (int)fminf(fmaxf(fy0, 0.0f), H), y1 = (int)fmaxf(fminf(fy1, H - 1.0f), -1.0f);
    for (int y = y0; y <= y1; ++y) {
        const float yn = 1.0f - ((float)y + 0.5f) * sy;
        const float v = (yn - cny) / rny;
        for (int x = x0; x <= x1; ++x) {
            const float xn = ((float)x + 0.5f) * sx - 1.0f;
            const float u = (xn - cnx) / rnx;
            const float r = sqrtf(u * u + v * v);                            // density.wgsl:80
            if (r > 1.0f) continue;                                          // density.wgsl:82-84
            const float w = render_exp(-10.0f * r * r);                      // density.wgsl:86-87
            const float d = base * w * rp.zoom_scale;                        // density.wgsl:94
            const uint64_t q = to_fixed(d);
            if (q) atomicAdd(&dens[(size_t)y * rp.width + x], (unsigned long long)q);
        }
    }
}

// fp16 bits of v * 2^-40, rounded to nearest even (inf above the largest half).
__device__ __forceinline__ uint32_t fixed40_to_f16(uint64_t v) {
    if (v == 0) return 0u;
    const int msb = 63 - __builtin_clzll(v);
    int e = msb - 40;
    uint64_t m, rem, half;
    if (e >= -14) {
        const int shift = msb - 10;  // >= 16
        m = v >> shift;
        rem = v & ((1ull << shift) - 1ull);
        half = 1ull << (shift - 1);
    } else {  // subnormal: multiples of 2^-24 = 2^16 units
        m = v >> 16;
        rem = v & 0xFFFFull;
        half = 0x8000ull;
        e = -15;
    }
    if (rem > half || (rem == half && (m & 1ull))) ++m;
    if (e == -15) return (uint32_t)m;  // 1024 is the smallest normal's bits
    if (m == 2048ull) {
        m = 1024ull;
        ++e;
    }
    if (e > 15) return 0x7C00u;
    return ((uint32_t)(e + 15) << 10) | (uint32_t)(m & 0x3FFull);
}

__device__ __forceinline__ double f16_value(uint32_t h) {
    const uint32_t ex = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    const double sgn = (h & 0x8000u) ? -1.0 : 1.0;
    if (ex == 0) return sgn * ldexp((double)mant, -24);
    if (ex == 31) return mant ? __builtin_nan("") : sgn * __builtin_huge_val();
    return sgn * ldexp((double)(mant | 0x400u), (int)ex - 25);
}

__device__ __forceinline__ double mixd(double a, double b, double t) { return a * (1.0 - t) + b * t; }

__device__ __forceinline__ uint32_t unorm8(double c) {
    const double q = rint(c * 255.0);
    return (uint32_t)(q < 0.0 ? 0.0 : (q > 255.0 ? 255.0 : q));
}

// tone_map.wgsl:66-84 for every fp16 density value h.
__global__ void tone_lut_kernel(uint32_t *__restrict__ lut) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= 65536u) return;
    const double density = f16_value(h);
    const double x = (double)8.0f * density;                  // exposure, tone_map.wgsl:72-73
    double mapped = x / (1.0 + x);                            // Reinhard, tone_map.wgsl:76
    if (!(mapped == mapped)) mapped = x > 0.0 ? 1.0 : 0.0;    // +inf / NaN density
    mapped = fmin(fmax(mapped, 0.0), 1.0);
    // gamma, tone_map.wgsl:79-80: the constant 1.0 / 2.2, rounded to binary32 where used
    const double t = pow(mapped, (double)(float)(1.0 / 2.2));
    // colormap, tone_map.wgsl:24-44 (single-precision constants)
    const double s1 = (double)0.33f, s2 = (double)0.66f, w3 = (double)0.34f;
    const double c1[3] = {(double)0.10f, (double)0.20f, (double)0.90f};
    const double c2[3] = {(double)0.20f, (double)0.90f, (double)0.90f};
    double rgb[3];
    for (int k = 0; k < 3; ++k) {
        if (t < s1)
            rgb[k] = mixd(0.0, c1[k], t / s1);
        else if (t < s2)
            rgb[k] = mixd(c1[k], c2[k], (t - s1) / s1);
        else
            rgb[k] = mixd(c2[k], 1.0, (t - s2) / w3);
    }
    lut[h] = unorm8(rgb[0]) | (unorm8(rgb[1]) << 8) | (unorm8(rgb[2]) << 16) | (255u << 24);
}

__global__ __launch_bounds__(256) void tone_map_kernel(const unsigned long long *__restrict__ dens,
                                                       const uint32_t *__restrict__ lut,
                                                       int width, int height,
                                                       uint16_t *__restrict__ d16,
                                                       uint32_t *__restrict__ rgba) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= width) return;
    const size_t texel = (size_t)(height - 1 - y) * width + x;
    const uint32_t h = fixed40_to_f16(dens[texel]);
    if (d16) d16[texel] = (uint16_t)h;
    rgba[(size_t)y * width + x] = lut[h];
}

}  // namespace

void launch_render(const float2 *pos, const float *radius, uint32_t n, const RenderParams &rp,
                   unsigned long long *dens, uint32_t *lut, bool build_lut, uint16_t *d16,
                   uint32_t *rgba, hipStream_t s) {
    const int64_t pixels = (int64_t)rp.width * rp.height;
    launch_clear_u32(reinterpret_cast<uint32_t *>(dens), 2 * pixels, s);
    if (build_lut) tone_lut_kernel<<<256, 256, 0, s>>>(lut);
    if (n > 0) splat_kernel<<<(n + 255u) / 256u, 256, 0, s>>>(pos, radius, n, rp, dens);
    dim3 grid((unsigned)((rp.width + 255) / 256), (unsigned)rp.height);
    tone_map_kernel<<<grid, 256, 0, s>>>(dens, lut, rp.width, rp.height, d16, rgba);
}

}  // namespace nbs
