// nb_ic.cpp -- seeded initial conditions (host).  See include/nbody/nbody_ic.h.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nbody/nbody_ic.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Counter-based stream: draw k of body i under `seed`.
struct Rng {
    uint64_t seed;
    uint64_t body;
    uint64_t k = 0;
    Rng(uint64_t s, uint64_t b) : seed(s), body(b) {}
    uint64_t next() { return mix64(mix64(seed ^ mix64(body)) + (k++) * 0xd1b54a32d192ed03ULL); }
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
    double normal() {
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    void direction(double& x, double& y, double& z) {
        double cz = 2.0 * uniform() - 1.0, ph = 2.0 * kPi * uniform();
        double s = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        x = s * std::cos(ph);
        y = s * std::sin(ph);
        z = cz;
    }
};

void normalize3(float& x, float& y, float& z) {
    float l = std::sqrt(x * x + y * y + z * z);
    if (l > 0.0f) {
        x /= l; y /= l; z /= l;
    }
}

// Plummer sphere by Aarseth, Henon & Wielen (1974): r from the inverse of the
// cumulative mass M(r) = r^3/(r^2+a^2)^{3/2}, speed q*v_esc with q drawn by
// rejection from q^2 (1-q^2)^{7/2}, isotropic directions; centre of mass moved
// to the origin in double precision.
void plummer(int n, uint64_t seed, double M, double a, double cutoff, double G, double* px, double* vx) {
    if (!(cutoff > 0.0 && cutoff <= 1.0)) cutoff = 0.999;
    for (int i = 0; i < n; ++i) {
        Rng r(seed, (uint64_t)i);
        double X1 = r.uniform() * cutoff;
        double rad = a / std::sqrt(std::pow(X1, -2.0 / 3.0) - 1.0);
        double dx, dy, dz;
        r.direction(dx, dy, dz);
        px[3 * i] = rad * dx; px[3 * i + 1] = rad * dy; px[3 * i + 2] = rad * dz;
        double q = 0.0;
        for (;;) {
            double X4 = r.uniform(), X5 = r.uniform();
            double g = X4 * X4 * std::pow(1.0 - X4 * X4, 3.5);
            if (0.1 * X5 < g) { q = X4; break; }
        }
        double ve = std::sqrt(2.0 * G * M) * std::pow(rad * rad + a * a, -0.25);
        double s = q * ve;
        r.direction(dx, dy, dz);
        vx[3 * i] = s * dx; vx[3 * i + 1] = s * dy; vx[3 * i + 2] = s * dz;
    }
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) { cp[k] += px[3 * i + k]; cv[k] += vx[3 * i + k]; }
    for (int k = 0; k < 3; ++k) { cp[k] /= n; cv[k] /= n; }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) { px[3 * i + k] -= cp[k]; vx[3 * i + k] -= cv[k]; }
}

}  // namespace

extern "C" {

void nb_ic_reference_default(nb_ic_reference_params* p) {
    if (!p) return;
    p->starting_speed_mul = 1.0f;
    p->starting_volume_radius = 1.5f;
    p->number_of_massive_objects = 0;
    p->massive_object_mass = 1.0f;
    p->pos_config = NB_UNIFORM_POS;
    p->vel_config = NB_FUNC_ZERO_VEL;
    p->mass_mean = 0.5f;
    p->mass_deviation = 0.25f;
}

void nb_ic_params_default(nb_ic_params* p, int32_t kind) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->total_mass = 1.0f;
    p->scale = 1.0f;
    p->mass_lo = 0.5f;
    p->mass_hi = 1.5f;
    p->cutoff = 0.999f;
    p->G = 1.0f;
    if (kind == NB_IC_COLLISION) {
        p->scale = 0.5f;
        p->center[0] = 2.0f;
        p->bulk_velocity[0] = -0.3f;
    }
}

// InitBodyAttributes (src/MyApp.cpp:331-495), seeded.
nb_status nb_ic_reference(const nb_ic_reference_params* p, int32_t n, uint64_t seed, float* pos, float* vel,
                          float* mass) {
    if (!p || n < 0 || (n > 0 && (!pos || !vel || !mass))) return NB_E_INVALID_ARG;
    const float R = p->starting_volume_radius, mul = p->starting_speed_mul;
    for (int i = 0; i < n; ++i) {
        Rng r(seed, (uint64_t)i);
        mass[i] = std::fabs((float)(p->mass_mean + p->mass_deviation * r.normal()));  // :340-343
    }
    // :345-355 keeps drawing indices in [0, K) until all of them are massive.
    for (int i = 0; i < p->number_of_massive_objects && i < n; ++i) mass[i] = p->massive_object_mass;
    if (p->pos_config == NB_SPHERE_POS) {
        // Fibonacci shell (:361-382): x = frac(i/phi), y = (i+0.5)/N,
        // lat = 2*2pi*x, lon = acos(1-2y).
        const float golden = 1.61803398874989484820f, two_pi = 6.28318530717958647692f;
        const float N = (float)n;
        float idx = 0.0f;
        for (int i = 0; i < n; ++i) {
            float tmp;
            float x = std::modf(idx / golden, &tmp);
            float y = (idx + 0.5f) / (N - 1.0f + 2.0f * 0.5f);
            float lat = 2.0f * two_pi * x;
            float lon = std::acos(1.0f - 2.0f * y);
            pos[3 * i] = R * std::sin(lat) * std::sin(lon);
            pos[3 * i + 1] = R * std::cos(lon);
            pos[3 * i + 2] = R * std::cos(lat) * std::sin(lon);
            idx += 1.0f;
        }
    } else {
        // iid Normal(0, R) per axis (:384-396)
        for (int i = 0; i < n; ++i) {
            Rng r(seed ^ 0x5bd1e995ULL, (uint64_t)i);
            for (int k = 0; k < 3; ++k) pos[3 * i + k] = (float)(R * r.normal());
        }
    }
    for (int i = 0; i < n; ++i) {
        float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        float vx = 0, vy = 0, vz = 0;
        switch (p->vel_config) {
            case NB_RANDOM_VEL: {  // normalize(Normal(0,0.1)^3) * mul (:421-437)
                Rng r(seed ^ 0x27d4eb2fULL, (uint64_t)i);
                vx = (float)(0.1 * r.normal()); vy = (float)(0.1 * r.normal()); vz = (float)(0.1 * r.normal());
                normalize3(vx, vy, vz);
                vx *= mul; vy *= mul; vz *= mul;
            } break;
            case NB_STARTING_OUT_VEL:  // (:438-449)
            case NB_STARTING_IN_VEL: { // (:450-459)
                vx = x; vy = y; vz = z;
                normalize3(vx, vy, vz);
                const float sgn = p->vel_config == NB_STARTING_OUT_VEL ? 1.0f : -1.0f;
                vx *= mul * sgn; vy *= mul * sgn; vz *= mul * sgn;
            } break;
            case NB_FUNC_ZERO_VEL:  // FLT_EPSILON per component (:460-469)
                vx = vy = vz = FLT_EPSILON;
                break;
            case NB_TANGENT_XZ_VEL: {  // cross(e_y, p), or cross(e_x, p) on the axis (:470-483)
                if (std::fabs(y) < FLT_EPSILON) {
                    vx = 0.0f; vy = -z; vz = y;
                } else {
                    vx = z; vy = 0.0f; vz = -x;
                }
                vx *= mul; vy *= mul; vz *= mul;
            } break;
            default:
                return NB_E_INVALID_ARG;
        }
        vel[3 * i] = vx; vel[3 * i + 1] = vy; vel[3 * i + 2] = vz;
    }
    return NB_OK;
}

nb_status nb_ic_generate(const nb_ic_params* p, int32_t n, uint64_t seed, float* pos, float* vel, float* mass) {
    if (!p || n < 0 || (n > 0 && (!pos || !vel || !mass))) return NB_E_INVALID_ARG;
    if (n == 0) return NB_OK;
    switch (p->kind) {
        case NB_IC_NORMAL:
            for (int i = 0; i < n; ++i) {
                Rng r(seed, (uint64_t)i);
                for (int k = 0; k < 3; ++k) pos[3 * i + k] = (float)(p->scale * r.normal()) + p->center[k];
                double u = r.uniform();
                mass[i] = (float)((p->mass_lo + (p->mass_hi - p->mass_lo) * u) / n);
                for (int k = 0; k < 3; ++k) vel[3 * i + k] = p->bulk_velocity[k];
            }
            return NB_OK;
        case NB_IC_UNIFORM_BALL:
            for (int i = 0; i < n; ++i) {
                Rng r(seed, (uint64_t)i);
                double rad = p->scale * std::cbrt(r.uniform());
                double dx, dy, dz;
                r.direction(dx, dy, dz);
                pos[3 * i] = (float)(rad * dx) + p->center[0];
                pos[3 * i + 1] = (float)(rad * dy) + p->center[1];
                pos[3 * i + 2] = (float)(rad * dz) + p->center[2];
                mass[i] = p->total_mass / (float)n;
                for (int k = 0; k < 3; ++k) vel[3 * i + k] = p->bulk_velocity[k];
            }
            return NB_OK;
        case NB_IC_PLUMMER: {
            std::vector<double> px(3 * (size_t)n), vx(3 * (size_t)n);
            plummer(n, seed, p->total_mass, p->scale, p->cutoff, p->G, px.data(), vx.data());
            for (int i = 0; i < n; ++i) {
                for (int k = 0; k < 3; ++k) {
                    pos[3 * i + k] = (float)(px[3 * i + k] + p->center[k]);
                    vel[3 * i + k] = (float)(vx[3 * i + k] + p->bulk_velocity[k]);
                }
                mass[i] = p->total_mass / (float)n;
            }
            return NB_OK;
        }
        case NB_IC_COLLISION: {
            // two equal Plummer spheres: bodies [0, n/2) at +center moving +v,
            // bodies [n/2, n) at -center moving -v; each sphere has total_mass
            const int n1 = n / 2, n2 = n - n1;
            for (int h = 0; h < 2; ++h) {
                const int m = h == 0 ? n1 : n2, off = h == 0 ? 0 : n1;
                if (m == 0) continue;
                const float sg = h == 0 ? 1.0f : -1.0f;
                std::vector<double> px(3 * (size_t)m), vx(3 * (size_t)m);
                plummer(m, mix64(seed + 1000003ULL * (h + 1)), p->total_mass, p->scale, p->cutoff, p->G, px.data(),
                        vx.data());
                for (int i = 0; i < m; ++i) {
                    for (int k = 0; k < 3; ++k) {
                        pos[3 * (off + i) + k] = (float)(px[3 * i + k] + sg * p->center[k]);
                        vel[3 * (off + i) + k] = (float)(vx[3 * i + k] + sg * p->bulk_velocity[k]);
                    }
                    mass[off + i] = p->total_mass / (float)m;
                }
            }
            return NB_OK;
        }
        default:
            return NB_E_INVALID_ARG;
    }
}

}  // extern "C"
