/*
 * nbody_ic.h -- seeded initial conditions (host only; no device needed).
 *
 * The reference draws its initial state unseeded inside
 * CMyApp::InitBodyAttributes (src/MyApp.cpp:331-504).  nb_ic_reference()
 * reproduces those modes with a counter-based generator, so the same seed
 * gives the same bytes on every machine.  The other kinds are the benchmark
 * distributions (Normal cloud, Plummer sphere, uniform ball, colliding pair).
 *
 * Random numbers: SplitMix64 finaliser over (seed, body, draw), i.e. a pure
 * function of its counters; Normal variates by Box-Muller in double.
 */
#ifndef NBODY_NBODY_IC_H
#define NBODY_NBODY_IC_H

#include <stdint.h>

#include "nbody/nbody.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SimulationConfig's initial-condition fields (src/SimulationConfig.h:85-100). */
typedef struct nb_ic_reference_params {
    float starting_speed_mul;        /* 1.0 */
    float starting_volume_radius;    /* 1.5 */
    int32_t number_of_massive_objects; /* 0 */
    float massive_object_mass;       /* 1.0 */
    int32_t pos_config;              /* NB_UNIFORM_POS */
    int32_t vel_config;              /* NB_FUNC_ZERO_VEL */
    float mass_mean;                 /* 0.5 */
    float mass_deviation;            /* 0.25 */
} nb_ic_reference_params;

enum { NB_IC_NORMAL = 0, NB_IC_PLUMMER = 1, NB_IC_UNIFORM_BALL = 2, NB_IC_COLLISION = 3 };

typedef struct nb_ic_params {
    int32_t kind;           /* NB_IC_* */
    float total_mass;       /* PLUMMER / BALL: total mass; COLLISION: mass of each sphere */
    float scale;            /* NORMAL: sigma; PLUMMER/COLLISION: scale radius a; BALL: radius */
    float mass_lo, mass_hi; /* NORMAL: m ~ U[mass_lo, mass_hi] / N */
    float cutoff;           /* PLUMMER: enclosed-mass fraction cutoff (e.g. 0.999) */
    float G;                /* PLUMMER: gravitational constant of the virial velocities */
    float center[3];        /* added to positions (COLLISION: sphere 1 at +center, sphere 2 at -center) */
    float bulk_velocity[3]; /* added to velocities (COLLISION: +v for sphere 1, -v for sphere 2) */
} nb_ic_params;

void nb_ic_reference_default(nb_ic_reference_params* p);
void nb_ic_params_default(nb_ic_params* p, int32_t kind);
/* Reference modes of InitBodyAttributes: masses |Normal(mean, dev)| with the
 * first K bodies massive, SPHERE_POS Fibonacci shell, UNIFORM_POS iid
 * Normal(0, R), five velocity modes. */
nb_status nb_ic_reference(const nb_ic_reference_params* p, int32_t n, uint64_t seed, float* pos, float* vel,
                          float* mass);
/* Benchmark distributions. */
nb_status nb_ic_generate(const nb_ic_params* p, int32_t n, uint64_t seed, float* pos, float* vel, float* mass);

#ifdef __cplusplus
}
#endif
#endif
