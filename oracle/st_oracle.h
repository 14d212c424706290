/*
 * st_oracle.h -- CPU restatement of the reference spin-tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (spin-tracking_amd/,
 * include/) may include, link or call this code; only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * This is plain C99 (gcc -O3 -ffp-contract=off, optional OpenMP).  Every
 * arithmetic expression keeps the operand order of the reference's
 * left-associative `coords` operator chains (src/double3.cpp:21-145), so
 * that +,-,*,/,sqrt round exactly as the reference does on x86-64.  The
 * libm transcendentals (cos, sin, acos, log, pow) are glibc's, as in the
 * reference's CPU build.
 *
 * Reference files followed (paths relative to the reference repository):
 *   include/options.h:7-61        options block, G_CONST, c2
 *   include/particle.h:52-59      rngState
 *   src/particle.cpp:24-119       sgn, RNG, variates
 *   src/particle.cpp:121-257      calc_next_collision_time
 *   src/particle.cpp:263-388      update_position_and_velocity
 *   src/particle.cpp:412-444      sanity_check
 *   src/particle.cpp:638-710      initParticlesCPU
 *   src/particle.cpp:712-837      runSimulationCPU
 *   src/integrator.cpp:20-88      sign/min_d/max_d, pulse, grad, interpolate,
 *                                 findCrossTerm, Bloch
 *   src/integrator.cpp:90-249     integrateDOP (+ include/coeff.h:6-176)
 *   src/integrator.cpp:251-337    integrateRK45 (+ include/coeff.h:178-218)
 *   src/integrator.cpp:339-723    types 2-5 (RKF45, quaternion forms, MagnusCFET)
 *   src/integrator.cpp:72-80, 724-789, src/floquet.cpp:25-162
 *                                 type 6: noise term, spectrum aggregator,
 *                                 integrateHamiltonian (the Floquet host math,
 *                                 src/floquet.cpp:188-301 and
 *                                 src/integrator.cpp:791-889, is in floquet.py)
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } or_coords;            /* include/double3.h:17-21 */

/* include/particle.h:52-59 -- 48 bytes on x86-64 */
typedef struct {
    uint64_t x, y, z, w;
    double   spare;
    bool     hasSpare;
} or_rng;

/* include/options.h:10-61 -- 440 bytes, byte-identical layout */
typedef struct {
    or_coords B0, E, L, yi, Gx, Gy, Gz;
    double m, t0, tf, rtol, atol, beta, uround, safe, fac1, fac2, hmax, hmin,
           h, T, sqrtKT_m, tc, gamma, V, a, w, swapStepSize, maxPosStep,
           ioutInt, diffuse;
    unsigned int nmax, seed;
    int integratorType, numParticles, numPerGPUBlock, iout;
    char dist;
    bool gas_coll, gravity, fixedStepSize, keepStepSize;
    or_coords noiseAmplitudes, noiseFrequencies;
} or_options;

/* Extension knobs that the reference hard-codes in its user-editable hooks
 * (src/integrator.cpp:40-42).  pulse_axis = 2 reproduces the reference hook
 * (0,0,a*cos(w*t)). */
typedef struct {
    int pulse_axis;      /* 0,1,2 = x,y,z */
} or_ext;

/* The reference's per-particle arrays (include/particle.h:147-170). */
typedef struct {
    or_coords *S, *v, *v_old, *pos, *pos_old;
    double *t, *t_old, *tf, *dt, *next_gas, *h;
    or_rng *state;
    uint64_t *n_bounce, *n_coll, *n_steps;
    unsigned int *partID;
    int *failure;
    bool *stop;
    char *coll_type, *wall_hit;
} or_arrays;

/* ---- RNG (src/particle.cpp:33-119, 646-663) ---- */
uint64_t or_xoshiro(or_rng *s);
double   or_uniform(or_rng *s);
double   or_normal(or_rng *s, double mean, double stdv);
double   or_maxboltz(or_rng *s, double sqrtkT_m);
double   or_exponential(or_rng *s, double tc);
void     or_seed(or_rng *s, unsigned int seed, unsigned int ipart);

/* ---- kinematics ---- */
void or_calc_next_collision_time(double t, double tf, or_coords v, or_coords pos,
        double *next_gas, double *dt, char *coll_type, uint64_t *n_bounce,
        uint64_t *n_coll, bool *finished, char *wall_hit, or_rng *state,
        const or_options *opt);
void or_update_position_and_velocity(double *t_old, double *t, double *dt,
        or_coords *pos_old, or_coords *pos, or_coords *v_old, or_coords *v,
        char *coll_type, char *wall_hit, or_rng *state, bool *stop,
        const or_options *opt);
void or_sanity_check(double *t_old, double *t, or_coords *pos_old, or_coords *pos,
        or_coords *v, or_coords *v_old, bool *stop, int *failure,
        const or_options *opt);

/* ---- spin ODE ---- */
void or_interpolate(double t, double t0, or_coords p_old, or_coords v_old,
        or_coords *p_out, or_coords *v_out, const or_options *opt);
or_coords or_cross_term(double t, const or_options *opt, const or_ext *ext,
        double t0, or_coords p_old, or_coords v_old);
or_coords or_bloch(double t, or_coords y, const or_options *opt, const or_ext *ext,
        double t0, or_coords p_old, or_coords v_old);
/* integratorType 0: integrateDOP, src/integrator.cpp:90-249 */
int or_integrate_dop(double t0, double tf, or_coords *y, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *h,
        uint64_t *n_acc, uint64_t *n_rej);
/* integratorType 1: integrateRK45, :251-337 */
int or_integrate_rk45(double t0, double tf, or_coords *y, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *h,
        uint64_t *n_acc, uint64_t *n_rej);
/* integratorType 2: integrateMagnusCFET, :620-723 */
int or_integrate_magnus(double t0, double tf, or_coords *y, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *h,
        uint64_t *n_acc, uint64_t *n_rej);
/* integratorType 3: integrateRK45Quaternion, :428-522 */
int or_integrate_rk45q(double t0, double tf, or_coords *y, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *h,
        uint64_t *n_acc, uint64_t *n_rej);
/* integratorType 4: integrateRKF45, :339-426 */
int or_integrate_rkf45(double t0, double tf, or_coords *y, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *h,
        uint64_t *n_acc, uint64_t *n_rej);
/* integratorType 5: integrateRKF45Quaternion, :524-618 */
int or_integrate_rkf45q(double t0, double tf, or_coords *y, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *h,
        uint64_t *n_acc, uint64_t *n_rej);

/* ---- spectral path, integratorType 6 (include/floquet.h:33-97,
 *      src/floquet.cpp:25-162, src/integrator.cpp:72-80, 724-789) ---- */
#define OR_NW 17   /* include/floquet.h:33-34 */
#define OR_NK 11
/* SpectrumAggregator's state (include/floquet.h:75-97); of covMat only
 * imag_diag ever changes (src/floquet.cpp:105-108) */
typedef struct {
    or_coords s1[OR_NW], s2[OR_NW], im[OR_NW];
    int n_samples;
} or_specagg;
/* SpectrumAggregator::w and ::dt, set once by initialize (:150-153) */
typedef struct {
    double w[OR_NW];
    double dt;
} or_speccfg;
or_coords or_noise_term(double t, const or_options *opt, double t0, or_coords p_old,
        or_coords v_old);
void or_spectrum_update(or_specagg *a, const or_speccfg *c, or_coords x);
int or_integrate_spectrum(double t0, double tf, or_specagg *agg, const or_speccfg *cfg,
        or_coords p_old, or_coords v_old, const or_options *opt, double h);
void or_covariance(const or_specagg *a, const or_speccfg *cfg, double out[OR_NW][3][3][2]);
void or_covariances(const or_specagg *agg, int n, const or_speccfg *cfg, double *out);
/* q = (w, x, y, z) in place */
int or_integrate_hamiltonian(double t0, double tf, double *q, const or_options *opt,
        const or_ext *ext, double h);
int or_sizeof_specagg(void);

/* ---- ensemble (initParticlesCPU / runSimulationCPU) ----
 * first/stride select the global particle indices first + j*stride for the
 * local slot j, so a deterministic subset of a large run can be replayed. */
void or_init_particles(const or_options *opt, const or_arrays *a, int n,
        unsigned int first, unsigned int stride);
/* Advances every particle to nextTOut.  stop_at (may be NULL) gives a
 * per-particle end time for runs with uneven durations: a particle takes
 * part in the interval (prevTOut, nextTOut] only if prevTOut < stop_at[j],
 * and its target is then min(nextTOut, stop_at[j]).
 * n_ode (may be NULL, 3 slots per particle) accumulates accepted trial
 * steps, rejected trial steps and segments. */
void or_run_simulation(const or_options *opt, const or_ext *ext, const or_arrays *a,
        int n, double nextTOut, const double *stop_at, double prevTOut,
        uint64_t *n_ode, int nthreads);

/* The same with the type-6 aggregators agg[n] (reset at the start of the
 * call) and their frequencies cfg. */
void or_run_simulation_spec(const or_options *opt, const or_ext *ext, const or_arrays *a,
        int n, double nextTOut, const double *stop_at, double prevTOut,
        uint64_t *n_ode, int nthreads, or_specagg *agg, const or_speccfg *cfg);

int or_sizeof_options(void);
int or_sizeof_rng(void);

#ifdef __cplusplus
}
#endif
#endif
