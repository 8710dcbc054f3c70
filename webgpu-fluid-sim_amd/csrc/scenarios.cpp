// scenarios.cpp -- seeded initial conditions (host code).
//
// The reference's setScenario() (src/simulation.ts:233-374) fills its
// arrays with the unseeded Math.random.  The three scenarios are restated
// here with every uniform draw taken from Philox-4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed
// by the seed and counted by (body, draw), so a given (seed, n) always
// yields the same bodies.  The benchmark workloads (uniform disk,
// exponential disk, colliding disks, Plummer sphere) use the same stream.
// All arithmetic is in double, as in the reference's JavaScript, and the
// results are stored as float32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/nbody_sim.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, hi0, lo0);
        mulhilo(0xCD9E8D57u, c2, hi1, lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Uniform doubles in (0, 1) for body i, block j (4 per block).
struct Stream {
    uint32_t key[2];
    uint32_t tag;
    void draw(uint64_t i, uint32_t j, double u[4]) const {
        const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), j, tag};
        uint32_t o[4];
        philox(ctr, key, o);
        for (int k = 0; k < 4; ++k) u[k] = ((double)o[k] + 0.5) * (1.0 / 4294967296.0);
    }
};

Stream make_stream(uint64_t seed, uint32_t tag) {
    Stream s;
    s.key[0] = (uint32_t)seed;
    s.key[1] = (uint32_t)(seed >> 32);
    s.tag = tag;
    return s;
}

// Box-Muller: two standard normals from two uniforms.
inline void gauss2(double u1, double u2, double &z0, double &z1) {
    const double r = std::sqrt(-2.0 * std::log(u1));
    z0 = r * std::cos(2.0 * kPi * u2);
    z1 = r * std::sin(2.0 * kPi * u2);
}

void set(float *mass, float *pos, float *vel, int64_t i, double m, double x, double y, double vx,
         double vy) {
    mass[i] = (float)m;
    pos[2 * i] = (float)x;
    pos[2 * i + 1] = (float)y;
    vel[2 * i] = (float)vx;
    vel[2 * i + 1] = (float)vy;
}

// Circular speeds sqrt(G*M(<r)/r) from the enclosed mass of the given
// bodies (positions relative to `cx, cy`), counted over strictly smaller radii.
std::vector<double> circular_speed(const std::vector<double> &r, const std::vector<double> &m,
                                   double g) {
    const size_t n = r.size();
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), (size_t)0);
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return r[a] < r[b]; });
    std::vector<double> vc(n, 0.0);
    double enclosed = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const size_t i = order[k];
        vc[i] = r[i] > 0.0 ? std::sqrt(g * enclosed / r[i]) : 0.0;
        enclosed += m[i];
    }
    return vc;
}

// Exponential disk of `count` bodies written at [first, first+count).
void exp_disk(const Stream &st, int64_t first, int64_t count, double scale, double total_mass,
              double mass_sigma, double disp, double g, double cx, double cy, double bvx,
              double bvy, float *mass, float *pos, float *vel) {
    std::vector<double> r(count), phi(count), m(count), z(2 * count);
    double msum = 0.0;
    for (int64_t k = 0; k < count; ++k) {
        double u[4], w[4];
        st.draw((uint64_t)(first + k), 0, u);
        st.draw((uint64_t)(first + k), 1, w);
        r[k] = -scale * std::log(u[0] * u[1]);  // Gamma(2, scale): p(r) ~ r exp(-r/scale)
        phi[k] = 2.0 * kPi * u[2];
        double g0, g1, g2, g3;
        gauss2(u[3], w[0], g0, g1);
        gauss2(w[1], w[2], g2, g3);
        (void)g1;
        m[k] = mass_sigma > 0.0 ? std::exp(mass_sigma * g0) : 1.0;
        z[2 * k] = g2;
        z[2 * k + 1] = g3;
        msum += m[k];
    }
    for (int64_t k = 0; k < count; ++k) m[k] *= total_mass / msum;
    std::vector<double> vc = circular_speed(r, m, g);
    for (int64_t k = 0; k < count; ++k) {
        const double c = std::cos(phi[k]), s = std::sin(phi[k]);
        const double vx = -vc[k] * s + disp * vc[k] * z[2 * k] + bvx;
        const double vy = vc[k] * c + disp * vc[k] * z[2 * k + 1] + bvy;
        set(mass, pos, vel, first + k, m[k], cx + r[k] * c, cy + r[k] * s, vx, vy);
    }
}

}  // namespace

extern "C" {

void nbody_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

int nbody_generate(int32_t scenario, int64_t n, uint64_t seed, const nbody_ic_params *params,
                   float *mass, float *pos, float *vel) {
    if (n < 0 || (n > 0 && (!mass || !pos || !vel))) return NBODY_ERR_INVALID_ARG;
    nbody_ic_params p{1.0f, 0.0f, 0.0f, 0.0f};
    if (params) p = *params;
    const double g = p.grav_constant;
    const Stream st = make_stream(seed, (uint32_t)scenario);
    const double N = (double)n;
    switch (scenario) {
        case NBODY_SCENARIO_DEFAULT: {  // simulation.ts:243-266
            const double radius = 5.0, speed_factor = 0.1;
            for (int64_t i = 0; i < n; ++i) {
                double u[4];
                st.draw((uint64_t)i, 0, u);
                const double angle = u[0] * 2.0 * kPi;
                const double r = radius * std::sqrt(u[1]);
                const double x = r * std::cos(angle), y = r * std::sin(angle);
                const double dist = std::sqrt(x * x + y * y);
                const double speed = speed_factor * std::sqrt(g * N / dist);
                set(mass, pos, vel, i, 1.0, x, y, -speed * (y / dist), speed * (x / dist));
            }
            return NBODY_OK;
        }
        case NBODY_SCENARIO_TWO_CLUSTERS: {  // simulation.ts:267-306
            const double offset = 15.0, radius = 3.0, speed_factor = 0.12;
            for (int64_t i = 0; i < n; ++i) {
                double u[4];
                st.draw((uint64_t)i, 0, u);
                const double angle = u[0] * 2.0 * kPi;
                const double r = radius * std::sqrt(u[1]);
                const double xb = r * std::cos(angle), yb = r * std::sin(angle);
                const bool first = (double)i < N / 2.0;
                const double x = first ? xb - offset : xb + offset;
                const double dist = std::sqrt(xb * xb + yb * yb);
                const double speed = speed_factor * std::sqrt(g * (N / 2.0) / dist);
                const double vx = first ? -speed * (yb / dist) : speed * (yb / dist);
                const double vy = first ? speed * (xb / dist) : -speed * (xb / dist);
                set(mass, pos, vel, i, 1.0, x, yb, vx, vy);
            }
            return NBODY_OK;
        }
        case NBODY_SCENARIO_PLANETARY_DISK: {  // simulation.ts:308-361
            const double disk_max = 15.0, disk_min = 0.1, central = 1000.0, planet = 10.0;
            const double body_mass = central / (100.0 * N);
            if (n > 0) set(mass, pos, vel, 0, central, 0.0, 0.0, 0.0, 0.0);
            if (n > 1) set(mass, pos, vel, 1, planet, 2.5, 0.0, 0.0, 0.8 * std::sqrt(g * central / 2.5));
            if (n > 2) set(mass, pos, vel, 2, planet, 12.0, 0.0, 0.0, std::sqrt(g * central / 12.5));
            for (int64_t i = 3; i < n; ++i) {
                double u[4];
                st.draw((uint64_t)i, 0, u);
                const double angle = u[0] * 2.0 * kPi;
                const double r = std::sqrt(u[1] * (disk_max * disk_max - disk_min * disk_min) +
                                           disk_min * disk_min);
                const double x = r * std::cos(angle), y = r * std::sin(angle);
                const double dist = std::sqrt(x * x + y * y);
                const double speed = 1.01 * std::sqrt(g * central / (2.0 + dist));
                set(mass, pos, vel, i, body_mass, x, y, -speed * (y / dist), speed * (x / dist));
            }
            return NBODY_OK;
        }
        case NBODY_SCENARIO_UNIFORM_DISK: {
            const double radius = p.scale > 0.0f ? p.scale : 1.0;
            for (int64_t i = 0; i < n; ++i) {
                double u[4];
                st.draw((uint64_t)i, 0, u);
                const double r = radius * std::sqrt(u[0]), phi = 2.0 * kPi * u[1];
                set(mass, pos, vel, i, 1.0 / N, r * std::cos(phi), r * std::sin(phi), 0.0, 0.0);
            }
            return NBODY_OK;
        }
        case NBODY_SCENARIO_EXP_DISK: {
            const double scale = p.scale > 0.0f ? p.scale : 0.2;
            exp_disk(st, 0, n, scale, 1.0, p.mass_sigma, p.dispersion, g, 0.0, 0.0, 0.0, 0.0, mass,
                     pos, vel);
            return NBODY_OK;
        }
        case NBODY_SCENARIO_COLLISION: {
            const double scale = p.scale > 0.0f ? p.scale : 0.15;
            const int64_t n0 = (n + 1) / 2;
            exp_disk(st, 0, n0, scale, (double)n0 / N, 0.0, p.dispersion, g, -0.6, 0.0, 0.0, 0.3,
                     mass, pos, vel);
            exp_disk(st, n0, n - n0, scale, (double)(n - n0) / N, 0.0, p.dispersion, g, 0.6, 0.0,
                     0.0, -0.3, mass, pos, vel);
            for (int64_t i = 0; i < n; ++i) mass[i] = (float)(1.0 / N);
            return NBODY_OK;
        }
        case NBODY_SCENARIO_PLUMMER: {
            // Projected Plummer profile: M(<R) = R^2 / (R^2 + a^2), so
            // R = a sqrt(u / (1 - u)); u is limited to 0.99 (R <= ~10a).
            const double a = p.scale > 0.0f ? p.scale : 0.1;
            for (int64_t i = 0; i < n; ++i) {
                double u[4];
                st.draw((uint64_t)i, 0, u);
                const double q = 0.99 * u[0];
                const double r = a * std::sqrt(q / (1.0 - q)), phi = 2.0 * kPi * u[1];
                set(mass, pos, vel, i, 1.0 / N, r * std::cos(phi), r * std::sin(phi), 0.0, 0.0);
            }
            return NBODY_OK;
        }
        default:
            return NBODY_ERR_INVALID_ARG;  // simulation.ts:364 throws for unknown scenarios
    }
}

int nbody_generate_radius(int32_t scenario, int64_t n, float *radius_multiplier) {
    if (n < 0 || (n > 0 && !radius_multiplier)) return NBODY_ERR_INVALID_ARG;
    if (scenario < NBODY_SCENARIO_DEFAULT || scenario > NBODY_SCENARIO_PLUMMER)
        return NBODY_ERR_INVALID_ARG;
    for (int64_t i = 0; i < n; ++i) radius_multiplier[i] = 1.0f;  // simulation.ts:240-241
    if (scenario == NBODY_SCENARIO_PLANETARY_DISK) {
        if (n > 0) radius_multiplier[0] = 12.0f;  // central body, simulation.ts:321
        if (n > 1) radius_multiplier[1] = 5.0f;   // planets, simulation.ts:330, 339
        if (n > 2) radius_multiplier[2] = 5.0f;
    }
    return NBODY_OK;
}

}  // extern "C"
