/*
 * st_oracle.c -- CPU restatement of the reference spin-tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see st_oracle.h).  Build flags: -O3
 * -ffp-contract=off -fno-fast-math; x86-64 SSE2 doubles, no excess precision.
 *
 * Each function cites the reference lines it restates.  Expressions are
 * written component by component in the order the reference's coords
 * operators evaluate them (src/double3.cpp:21-145): `a + b*c + d` is
 * (a + (b*c)) + d, `s * v * w` is ((s*v)*w), and so on.
 */
#include "st_oracle.h"
/* the field hooks pulse()/grad() (src/integrator.cpp:40-46), shared with the
 * HIP kernels so both evaluate the same, possibly user-edited, fields */
#include "../include/spintrack_fields.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* include/options.h:7-8 */
static const double G_CONST = -9.81;
static const double C2 = 299792458.0 * 299792458.0;

_Static_assert(sizeof(or_options) == 440, "options block must be 440 bytes");
_Static_assert(offsetof(or_options, nmax) == 360, "options layout");
_Static_assert(offsetof(or_options, dist) == 384, "options layout");
_Static_assert(offsetof(or_options, noiseAmplitudes) == 392, "options layout");
_Static_assert(sizeof(or_rng) == 48, "rngState must be 48 bytes");

int or_sizeof_options(void) { return (int)sizeof(or_options); }
int or_sizeof_rng(void) { return (int)sizeof(or_rng); }

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */

/* src/particle.cpp:24-31: sgn(0) = +1 */
static inline double psgn(double v) { return (v < 0) ? -1.0 : 1.0; }
/* std::min / std::max as the reference calls them (b<a ? b : a), (a<b ? b : a) */
static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }
/* src/integrator.cpp:20-33 */
static inline double sign_d(double a, double b) { return (b < 0.0) ? -fabs(a) : fabs(a); }
static inline double min_d(double a, double b) { return (a < b) ? a : b; }
static inline double max_d(double a, double b) { return (a > b) ? a : b; }

/* src/double3.cpp:85-91 */
static inline or_coords cross3(or_coords a, or_coords b) {
    or_coords o;
    o.x = a.y * b.z - a.z * b.y;
    o.y = a.z * b.x - a.x * b.z;
    o.z = a.x * b.y - a.y * b.x;
    return o;
}
/* src/double3.cpp:93-95 */
static inline double dot3(or_coords a, or_coords b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* src/double3.cpp:101-103 */
static inline double len3(or_coords a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

/* ------------------------------------------------------------------ */
/* RNG: src/particle.cpp:33-119                                         */
/* ------------------------------------------------------------------ */

static inline uint64_t rol64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* src/particle.cpp:59-70 -- the reference's own xoshiro variant: output x+z,
 * and w is *replaced* by rotl(z, 45). */
uint64_t or_xoshiro(or_rng *s) {
    const uint64_t result = s->x + s->z;
    const uint64_t t = s->y << 17;
    s->z ^= s->x;
    s->w ^= s->y;
    s->y ^= s->z;
    s->x ^= s->w;
    s->z ^= t;
    s->w = rol64(s->z, 45);
    return result;
}

/* src/particle.cpp:72-83 */
double or_uniform(or_rng *s) {
    uint64_t r = or_xoshiro(s);
    return (double)(r >> 11) * 0x1.0p-53;
}

/* src/particle.cpp:85-87 */
static inline double uniform_lh(or_rng *s, double low, double high) {
    return or_uniform(s) * (high - low) + low;
}

/* src/particle.cpp:89-107 -- Marsaglia polar method with a cached spare */
double or_normal(or_rng *s, double mean, double stdv) {
    if (s->hasSpare) {
        s->hasSpare = false;
        return s->spare * stdv + mean;
    }
    double ts, tu, tv;
    do {
        tu = uniform_lh(s, -1.0, 1.0);
        tv = uniform_lh(s, -1.0, 1.0);
        ts = tu * tu + tv * tv;
    } while (ts >= 1.0 || ts == 0.0);
    ts = sqrt(-2.0 * log(ts) / ts);
    s->spare = tv * ts;
    s->hasSpare = true;
    return mean + stdv * tu * ts;
}

/* src/particle.cpp:109-111 */
double or_maxboltz(or_rng *s, double sqrtkT_m) { return or_normal(s, 0.0, 1.0) * sqrtkT_m; }
/* src/particle.cpp:113-115 */
static inline double unif02pi(or_rng *s) { return or_uniform(s) * 2.0 * M_PI; }
/* src/particle.cpp:117-119 */
double or_exponential(or_rng *s, double tc) { return -tc * log(1.0 - or_uniform(s)); }

/* src/particle.cpp:646-663: splitmix64 of (seed + index), computed in 32-bit
 * unsigned arithmetic and widened, split into 32-bit halves; then three
 * discarded uniforms. */
void or_seed(or_rng *s, unsigned int seed, unsigned int ipart) {
    uint64_t st = (uint64_t)(unsigned int)(seed + ipart);
    uint64_t r = (st += 0x9E3779B97f4A7C15ULL);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
    r = r ^ (r >> 31);
    s->x = (uint32_t)r;
    s->y = (uint32_t)(r >> 32);
    r = (st += 0x9E3779B97f4A7C15ULL);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
    r = r ^ (r >> 31);
    s->z = (uint32_t)r;
    s->w = (uint32_t)(r >> 32);
    s->spare = 0.0;
    s->hasSpare = false;
    or_uniform(s);
    or_uniform(s);
    or_uniform(s);
}

/* ------------------------------------------------------------------ */
/* Kinematics: src/particle.cpp:121-444                                 */
/* ------------------------------------------------------------------ */

/* src/particle.cpp:121-257 */
void or_calc_next_collision_time(double t, double tf, or_coords v, or_coords pos,
        double *next_gas, double *dt, char *coll_type, uint64_t *n_bounce,
        uint64_t *n_coll, bool *finished, char *wall_hit, or_rng *state,
        const or_options *opt) {
    double dx, dy, dz, dtx, dty, dtz;
    if (opt->gravity) {
        dx = psgn(v.x) * opt->L.x / 2.0 - pos.x;
        dz = psgn(v.z) * opt->L.z / 2.0 - pos.z;
        dtx = dx / v.x;
        dtz = dz / v.z;
        double y2 = v.y * v.y;
        if (psgn(v.y) <= 0.0) {
            dy = pos.y + opt->L.y * 0.5;
            double sqr = sqrt(-2.0 * G_CONST * dy + y2);
            double temp1 = -(sqr + v.y) / G_CONST;
            double temp2 = (sqr - v.y) / G_CONST;
            dty = std_min(fabs(temp1), fabs(temp2));
        } else {
            double maxHeight = -0.5 * y2 / G_CONST + pos.y;
            if (maxHeight < 0.5 * opt->L.y) {
                double t1 = -v.y / G_CONST;
                double topOfFlight = pos.y + v.y * t1 + 0.5 * G_CONST * t1 * t1;
                double t2 = sqrt((-0.5 * opt->L.y - topOfFlight) * 2.0 / G_CONST);
                dty = t1 + t2;
            } else {
                double sqr = sqrt(y2 - 2.0 * G_CONST * (pos.y - (0.5 * opt->L.y)));
                double temp1 = (-v.y + sqr) / G_CONST;
                double temp2 = (-v.y - sqr) / G_CONST;
                dty = std_min(fabs(temp1), fabs(temp2));
            }
        }
        /* an else-if chain: at most one fix-up per call (:152-157) */
        if (dtx < 1e-16 || isnan(dtx))
            dtx = 1e6;
        else if (dty < 1e-16 || isnan(dty))
            dty = 1e6;
        else if (dtz < 1e-16 || isnan(dtz))
            dtz = 1e6;
    } else {
        dx = psgn(v.x) * opt->L.x / 2.0 - pos.x;
        dy = psgn(v.y) * opt->L.y / 2.0 - pos.y;
        dz = psgn(v.z) * opt->L.z / 2.0 - pos.z;
        dtx = dx / v.x;
        dty = dy / v.y;
        dtz = dz / v.z;
        if (dtx < 1e-16)
            dtx = 1e6;
        else if (dty < 1e-16)
            dty = 1e6;
        else if (dtz < 1e-16)
            dtz = 1e6;
    }
    /* The reference leaves tbounce uninitialised when no branch matches (only
     * possible with NaN hit times); NaN is the deterministic stand-in. */
    double tbounce = NAN;
    if (dtx < dty && dtx < dtz) {
        tbounce = dtx;
        *wall_hit = 'x';
    } else if (dty < dtx && dty < dtz) {
        tbounce = dty;
        *wall_hit = 'y';
    } else if (dtz < dtx && dtz < dty) {
        tbounce = dtz;
        *wall_hit = 'z';
    } else {
        if (dtx == dty && dtx == dtz) {
            tbounce = dtx;
            *wall_hit = 'a';
        } else {
            if (dtx == dty) {
                tbounce = dtx;
                *wall_hit = 'b';
            } else if (dtx == dtz) {
                tbounce = dtx;
                *wall_hit = 'c';
            } else if (dty == dtz) {
                tbounce = dty;
                *wall_hit = 'd';
            }
        }
    }
    double timeToNextGas = *next_gas - t;
    int nextCol;
    double timeToNextCol;
    if (timeToNextGas < tbounce) {
        nextCol = 0;
        timeToNextCol = timeToNextGas;
    } else if (timeToNextGas == tbounce) {
        *next_gas += or_exponential(state, opt->tc);
        timeToNextCol = timeToNextGas;
        nextCol = 1;
    } else {
        nextCol = 1;
        timeToNextCol = tbounce;
    }
    int nextNothing;
    double nextNoCol;
    if (tf - t <= opt->maxPosStep) {
        nextNothing = 0;
        nextNoCol = tf - t;
    } else {
        nextNothing = 1;
        nextNoCol = opt->maxPosStep;
    }
    if (timeToNextCol <= nextNoCol) {
        *dt = timeToNextCol;
        if (nextCol == 0) {
            *coll_type = 'G';
            *n_coll += 1;
            *next_gas += or_exponential(state, opt->tc);
        } else {
            *n_bounce += 1;
            *coll_type = 'W';
        }
        if (timeToNextCol == nextNoCol) {
            if (nextNothing == 0) *finished = true;
        }
    } else {
        *coll_type = 'N';
        *dt = nextNoCol;
        if (nextNothing == 0) *finished = true;
    }
}

static inline void flip_velocity(or_coords *v, char wall_hit) {
    /* src/particle.cpp:329-371 */
    if (wall_hit == 'x')
        v->x *= -1.0;
    else if (wall_hit == 'y')
        v->y *= -1.0;
    else if (wall_hit == 'z')
        v->z *= -1.0;
    else if (wall_hit == 'a') {
        v->x *= -1.0;
        v->y *= -1.0;
        v->z *= -1.0;
    } else if (wall_hit == 'b') {
        v->x *= -1.0;
        v->y *= -1.0;
    } else if (wall_hit == 'c') {
        v->x *= -1.0;
        v->z *= -1.0;
    } else if (wall_hit == 'd') {
        v->y *= -1.0;
        v->z *= -1.0;
    }
}

/* src/particle.cpp:263-388 */
void or_update_position_and_velocity(double *t_old, double *t, double *dt,
        or_coords *pos_old, or_coords *pos, or_coords *v_old, or_coords *v,
        char *coll_type, char *wall_hit, or_rng *state, bool *stop,
        const or_options *opt) {
    *v_old = *v;
    *t_old = *t;
    *t += *dt;
    *pos_old = *pos;
    *v_old = *v;
    or_coords a = {0.0, 0.0, 0.0};
    if (opt->gravity) a.y = G_CONST;
    const double d = *dt;
    /* pos = pos_old + v*dt + 0.5*a*dt*dt ; v = v + a*dt */
    pos->x = (pos_old->x + v->x * d) + ((0.5 * a.x) * d) * d;
    pos->y = (pos_old->y + v->y * d) + ((0.5 * a.y) * d) * d;
    pos->z = (pos_old->z + v->z * d) + ((0.5 * a.z) * d) * d;
    v->x = v->x + a.x * d;
    v->y = v->y + a.y * d;
    v->z = v->z + a.z * d;
    const double Vel = len3(*v);
    const char ct = *coll_type, wh = *wall_hit;
    if (ct == 'N') {
        /* free flight, nothing to do */
    } else if (ct == 'W') {
        if (wh == 'x') {
            pos->x = psgn(pos->x) * opt->L.x * 0.5;
        } else if (wh == 'y') {
            pos->y = psgn(pos->y) * opt->L.y * 0.5;
        } else if (wh == 'z') {
            pos->z = psgn(pos->z) * opt->L.z * 0.5;
        } else if (wh == 'a') {
            pos->x = psgn(pos->x) * opt->L.x * 0.5;
            pos->y = psgn(pos->y) * opt->L.y * 0.5;
            pos->z = psgn(pos->z) * opt->L.z * 0.5;
        } else if (wh == 'b') {
            pos->x = psgn(pos->x) * opt->L.x * 0.5;
            pos->y = psgn(pos->y) * opt->L.y * 0.5;
        } else if (wh == 'c') {
            pos->x = psgn(pos->x) * opt->L.x * 0.5;
            pos->z = psgn(pos->z) * opt->L.z * 0.5;
        } else if (wh == 'd') {
            pos->y = psgn(pos->y) * opt->L.y * 0.5;
            pos->z = psgn(pos->z) * opt->L.z * 0.5;
        }
        if (opt->diffuse > FLT_MIN) {
            bool diffuse = or_uniform(state) < opt->diffuse;
            if (diffuse) {
                double phi = acos(sqrt(or_uniform(state)));
                double theta = unif02pi(state);
                if (wh == 'x') {
                    v->x = -1.0 * psgn(v->x) * Vel * cos(phi);
                    v->y = -Vel * sin(phi) * cos(theta);
                    v->z = Vel * sin(phi) * sin(theta);
                } else if (wh == 'y') {
                    v->x = Vel * sin(phi) * cos(theta);
                    v->y = -1.0 * psgn(v->y) * Vel * cos(phi);
                    v->z = Vel * sin(phi) * sin(theta);
                } else if (wh == 'z') {
                    v->x = Vel * sin(phi) * cos(theta);
                    v->y = Vel * sin(phi) * sin(theta);
                    v->z = -1.0 * psgn(v->z) * Vel * cos(phi);
                } else {
                    *stop = true;
                }
            } else {
                flip_velocity(v, wh);
            }
        } else {
            flip_velocity(v, wh);
        }
    } else if (ct == 'G') {
        if (opt->dist == 'M') {
            v->x = or_maxboltz(state, opt->sqrtKT_m);
            v->y = or_maxboltz(state, opt->sqrtKT_m);
            v->z = or_maxboltz(state, opt->sqrtKT_m);
        } else if (opt->dist == 'C') {
            or_coords vec;
            vec.x = or_normal(state, 0.0, 1.0);
            vec.y = or_normal(state, 0.0, 1.0);
            vec.z = or_normal(state, 0.0, 1.0);
            double vn = len3(vec);
            v->x = (Vel * vec.x) / vn;
            v->y = (Vel * vec.y) / vn;
            v->z = (Vel * vec.z) / vn;
        }
    }
}

/* src/particle.cpp:412-444 */
void or_sanity_check(double *t_old, double *t, or_coords *pos_old, or_coords *pos,
        or_coords *v, or_coords *v_old, bool *stop, int *failure,
        const or_options *opt) {
    const double tol = 1.0e-15;
    if (fabs(pos->x) > opt->L.x / 2.0 + tol) { *stop = true; *failure = 1; return; }
    if (fabs(pos->y) > opt->L.y / 2.0 + tol) { *stop = true; *failure = 1; return; }
    if (fabs(pos->z) > opt->L.z / 2.0 + tol) { *stop = true; *failure = 1; return; }
    if (isinf(*t_old) || isinf(*t)) { *stop = true; *failure = 2; return; }
    if (isnan(pos_old->x) || isnan(pos_old->y) || isnan(pos_old->z) ||
        isnan(pos->x) || isnan(pos->y) || isnan(pos->z)) {
        *stop = true; *failure = 3; return;
    }
    if (isnan(v->x) || isnan(v->y) || isnan(v->z) ||
        isnan(v_old->x) || isnan(v_old->y) || isnan(v_old->z)) {
        *failure = 4; *stop = true; return;
    }
}

/* ------------------------------------------------------------------ */
/* Bloch RHS: src/integrator.cpp:40-88                                  */
/* ------------------------------------------------------------------ */

/* src/integrator.cpp:48-61 */
void or_interpolate(double t, double t0, or_coords p_old, or_coords v_old,
        or_coords *p_out, or_coords *v_out, const or_options *opt) {
    if (opt->gravity) {
        const or_coords a = {0.0, G_CONST, 0.0};
        const double d = t - t0;
        p_out->x = (p_old.x + v_old.x * d) + ((0.5 * a.x) * d) * d;
        p_out->y = (p_old.y + v_old.y * d) + ((0.5 * a.y) * d) * d;
        p_out->z = (p_old.z + v_old.z * d) + ((0.5 * a.z) * d) * d;
        v_out->x = v_old.x + a.x * d;
        v_out->y = v_old.y + a.y * d;
        v_out->z = v_old.z + a.z * d;
    } else {
        const double d = t - t0;
        p_out->x = p_old.x + v_old.x * d;
        p_out->y = p_old.y + v_old.y * d;
        p_out->z = p_old.z + v_old.z * d;
        *v_out = v_old;
    }
}

/* src/integrator.cpp:40-42 (pulse), :44-46 (grad), :63-70 (findCrossTerm):
 * gamma * (pulse + B0 + (1/c^2) v x E + G.r) */
/* grad(pos, opt) through the hook: {dot(p, Gx), dot(p, Gy), dot(p, Gz)} for
 * the default hook (src/integrator.cpp:44-46) */
static or_coords field_grad(or_coords p, const or_options *opt) {
    const double r[3] = {p.x, p.y, p.z};
    const double G[9] = {opt->Gx.x, opt->Gx.y, opt->Gx.z, opt->Gy.x, opt->Gy.y, opt->Gy.z,
                         opt->Gz.x, opt->Gz.y, opt->Gz.z};
    double B[3];
    st_field_grad(r, G, B);
    or_coords g = {B[0], B[1], B[2]};
    return g;
}

or_coords or_cross_term(double t, const or_options *opt, const or_ext *ext,
        double t0, or_coords p_old, or_coords v_old) {
    or_coords p, v;
    or_interpolate(t, t0, p_old, v_old, &p, &v, opt);
    const or_coords G = field_grad(p, opt);
    /* pulse(t, a, w), evaluated even when a == 0 as the reference does
     * (src/integrator.cpp:67) */
    double P3[3];
    st_field_pulse(t, opt->a, opt->w, ext ? ext->pulse_axis : 2, P3);
    const or_coords pl = {P3[0], P3[1], P3[2]};
    const or_coords vxE = cross3(v, opt->E);
    const double ic2 = 1.0 / C2;
    or_coords B;
    B.x = ((pl.x + opt->B0.x) + vxE.x * ic2) + G.x;
    B.y = ((pl.y + opt->B0.y) + vxE.y * ic2) + G.y;
    B.z = ((pl.z + opt->B0.z) + vxE.z * ic2) + G.z;
    or_coords w;
    w.x = B.x * opt->gamma;
    w.y = B.y * opt->gamma;
    w.z = B.z * opt->gamma;
    return w;
}

/* src/integrator.cpp:83-88 (Bloch): f = y x findCrossTerm */
or_coords or_bloch(double t, or_coords y, const or_options *opt, const or_ext *ext,
        double t0, or_coords p_old, or_coords v_old) {
    return cross3(y, or_cross_term(t, opt, ext, t0, p_old, v_old));
}

/* ------------------------------------------------------------------ */
/* DOP853: src/integrator.cpp:90-249, tableau include/coeff.h:6-176     */
/* (Hairer & Wanner's published DOP853 coefficients)                    */
/* ------------------------------------------------------------------ */

#define Dc2 0.526001519587677318785587544488E-01
#define Dc3 0.789002279381515978178381316732E-01
#define Dc4 0.118350341907227396726757197510E+00
#define Dc5 0.281649658092772603273242802490E+00
#define Dc6 0.333333333333333333333333333333E+00
#define Dc7 0.25E+00
#define Dc8 0.307692307692307692307692307692E+00
#define Dc9 0.651282051282051282051282051282E+00
#define Dc10 0.6E+00
#define Dc11 0.857142857142857142857142857142E+00
#define Db1 5.42937341165687622380535766363E-2
#define Db6 4.45031289275240888144113950566E0
#define Db7 1.89151789931450038304281599044E0
#define Db8 -5.8012039600105847814672114227E0
#define Db9 3.1116436695781989440891606237E-1
#define Db10 -1.52160949662516078556178806805E-1
#define Db11 2.01365400804030348374776537501E-1
#define Db12 4.47106157277725905176885569043E-2
#define Dbhh1 0.244094488188976377952755905512E+00
#define Dbhh2 0.733846688281611857341361741547E+00
#define Dbhh3 0.220588235294117647058823529412E-01
#define Der1 0.1312004499419488073250102996E-01
#define Der6 -0.1225156446376204440720569753E+01
#define Der7 -0.4957589496572501915214079952E+00
#define Der8 0.1664377182454986536961530415E+01
#define Der9 -0.3503288487499736816886487290E+00
#define Der10 0.3341791187130174790297318841E+00
#define Der11 0.8192320648511571246570742613E-01
#define Der12 -0.2235530786388629525884427845E-01
#define Da21 5.26001519587677318785587544488E-2
#define Da31 1.97250569845378994544595329183E-2
#define Da32 5.91751709536136983633785987549E-2
#define Da41 2.95875854768068491816892993775E-2
#define Da43 8.87627564304205475450678981324E-2
#define Da51 2.41365134159266685502369798665E-1
#define Da53 -8.84549479328286085344864962717E-1
#define Da54 9.24834003261792003115737966543E-1
#define Da61 3.7037037037037037037037037037E-2
#define Da64 1.70828608729473871279604482173E-1
#define Da65 1.25467687566822425016691814123E-1
#define Da71 3.7109375E-2
#define Da74 1.70252211019544039314978060272E-1
#define Da75 6.02165389804559606850219397283E-2
#define Da76 -1.7578125E-2
#define Da81 3.70920001185047927108779319836E-2
#define Da84 1.70383925712239993810214054705E-1
#define Da85 1.07262030446373284651809199168E-1
#define Da86 -1.53194377486244017527936158236E-2
#define Da87 8.27378916381402288758473766002E-3
#define Da91 6.24110958716075717114429577812E-1
#define Da94 -3.36089262944694129406857109825E0
#define Da95 -8.68219346841726006818189891453E-1
#define Da96 2.75920996994467083049415600797E1
#define Da97 2.01540675504778934086186788979E1
#define Da98 -4.34898841810699588477366255144E1
#define Da101 4.77662536438264365890433908527E-1
#define Da104 -2.48811461997166764192642586468E0
#define Da105 -5.90290826836842996371446475743E-1
#define Da106 2.12300514481811942347288949897E1
#define Da107 1.52792336328824235832596922938E1
#define Da108 -3.32882109689848629194453265587E1
#define Da109 -2.03312017085086261358222928593E-2
#define Da111 -9.3714243008598732571704021658E-1
#define Da114 5.18637242884406370830023853209E0
#define Da115 1.09143734899672957818500254654E0
#define Da116 -8.14978701074692612513997267357E0
#define Da117 -1.85200656599969598641566180701E1
#define Da118 2.27394870993505042818970056734E1
#define Da119 2.49360555267965238987089396762E0
#define Da1110 -3.0467644718982195003823669022E0
#define Da121 2.27331014751653820792359768449E0
#define Da124 -1.05344954667372501984066689879E1
#define Da125 -2.00087205822486249909675718444E0
#define Da126 -1.79589318631187989172765950534E1
#define Da127 2.79488845294199600508499808837E1
#define Da128 -2.85899827713502369474065508674E0
#define Da129 -8.87285693353062954433549289258E0
#define Da1210 1.23605671757943030647266201528E1
#define Da1211 6.43392746015763530355970484046E-1

#define BLOCH(tt, yy) or_bloch((tt), (yy), opt, ext, t0, p_old, v_old)

int or_integrate_dop(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej) {
    or_coords y = *yio;
    double h = *hio;
    or_coords yy1, k1, k2, k3, k4, k5, k6, k7, k8, k9, k10;
    int reject, last;
    double facold, expo1, fac, facc1, facc2, fac11, posneg, xph, err2, deno, err, hnew;
    unsigned int nstep = 0;
    double x = t0;
    const double xf = tf;
    const int n = 3;
    facold = 1.0E-4;
    expo1 = 1.0 / 8.0 - opt->beta * 0.2;
    facc1 = 1.0 / opt->fac1;
    facc2 = 1.0 / opt->fac2;
    posneg = sign_d(1.0, tf - t0);
    last = 0;
    k1 = BLOCH(x, y);
    reject = 0;
    int ret;
    for (;;) {
        if (nstep > opt->nmax) { ret = -1; break; }
        if ((x + 1.01 * h - xf) * posneg > 0.0) {
            h = xf - x;
            last = 1;
        }
        if (h < opt->hmin && last != 1) h = opt->hmin;
        nstep++;

#define ST(dst, expr_x, expr_y, expr_z) do { (dst).x = (expr_x); (dst).y = (expr_y); (dst).z = (expr_z); } while (0)
        /* yy1 = y + h * a21 * k1   ==  y + k1*(h*a21) */
        { const double ha = h * Da21; ST(yy1, y.x + k1.x * ha, y.y + k1.y * ha, y.z + k1.z * ha); }
        k2 = BLOCH(x + Dc2 * h, yy1);
#define C2T(c, A, a1, B, b1) (A.c * a1 + B.c * b1)
        ST(yy1, y.x + h * (Da31 * k1.x + Da32 * k2.x),
                y.y + h * (Da31 * k1.y + Da32 * k2.y),
                y.z + h * (Da31 * k1.z + Da32 * k2.z));
        k3 = BLOCH(x + Dc3 * h, yy1);
        ST(yy1, y.x + h * (Da41 * k1.x + Da43 * k3.x),
                y.y + h * (Da41 * k1.y + Da43 * k3.y),
                y.z + h * (Da41 * k1.z + Da43 * k3.z));
        k4 = BLOCH(x + Dc4 * h, yy1);
        ST(yy1, y.x + h * (Da51 * k1.x + Da53 * k3.x + Da54 * k4.x),
                y.y + h * (Da51 * k1.y + Da53 * k3.y + Da54 * k4.y),
                y.z + h * (Da51 * k1.z + Da53 * k3.z + Da54 * k4.z));
        k5 = BLOCH(x + Dc5 * h, yy1);
        ST(yy1, y.x + h * (Da61 * k1.x + Da64 * k4.x + Da65 * k5.x),
                y.y + h * (Da61 * k1.y + Da64 * k4.y + Da65 * k5.y),
                y.z + h * (Da61 * k1.z + Da64 * k4.z + Da65 * k5.z));
        k6 = BLOCH(x + Dc6 * h, yy1);
        ST(yy1, y.x + h * (Da71 * k1.x + Da74 * k4.x + Da75 * k5.x + Da76 * k6.x),
                y.y + h * (Da71 * k1.y + Da74 * k4.y + Da75 * k5.y + Da76 * k6.y),
                y.z + h * (Da71 * k1.z + Da74 * k4.z + Da75 * k5.z + Da76 * k6.z));
        k7 = BLOCH(x + Dc7 * h, yy1);
        ST(yy1, y.x + h * (Da81 * k1.x + Da84 * k4.x + Da85 * k5.x + Da86 * k6.x + Da87 * k7.x),
                y.y + h * (Da81 * k1.y + Da84 * k4.y + Da85 * k5.y + Da86 * k6.y + Da87 * k7.y),
                y.z + h * (Da81 * k1.z + Da84 * k4.z + Da85 * k5.z + Da86 * k6.z + Da87 * k7.z));
        k8 = BLOCH(x + Dc8 * h, yy1);
        ST(yy1, y.x + h * (Da91 * k1.x + Da94 * k4.x + Da95 * k5.x + Da96 * k6.x + Da97 * k7.x + Da98 * k8.x),
                y.y + h * (Da91 * k1.y + Da94 * k4.y + Da95 * k5.y + Da96 * k6.y + Da97 * k7.y + Da98 * k8.y),
                y.z + h * (Da91 * k1.z + Da94 * k4.z + Da95 * k5.z + Da96 * k6.z + Da97 * k7.z + Da98 * k8.z));
        k9 = BLOCH(x + Dc9 * h, yy1);
        ST(yy1, y.x + h * (Da101 * k1.x + Da104 * k4.x + Da105 * k5.x + Da106 * k6.x + Da107 * k7.x + Da108 * k8.x + Da109 * k9.x),
                y.y + h * (Da101 * k1.y + Da104 * k4.y + Da105 * k5.y + Da106 * k6.y + Da107 * k7.y + Da108 * k8.y + Da109 * k9.y),
                y.z + h * (Da101 * k1.z + Da104 * k4.z + Da105 * k5.z + Da106 * k6.z + Da107 * k7.z + Da108 * k8.z + Da109 * k9.z));
        k10 = BLOCH(x + Dc10 * h, yy1);
        ST(yy1, y.x + h * (Da111 * k1.x + Da114 * k4.x + Da115 * k5.x + Da116 * k6.x + Da117 * k7.x + Da118 * k8.x + Da119 * k9.x + Da1110 * k10.x),
                y.y + h * (Da111 * k1.y + Da114 * k4.y + Da115 * k5.y + Da116 * k6.y + Da117 * k7.y + Da118 * k8.y + Da119 * k9.y + Da1110 * k10.y),
                y.z + h * (Da111 * k1.z + Da114 * k4.z + Da115 * k5.z + Da116 * k6.z + Da117 * k7.z + Da118 * k8.z + Da119 * k9.z + Da1110 * k10.z));
        k2 = BLOCH(x + Dc11 * h, yy1);
        xph = x + h;
        ST(yy1, y.x + h * (Da121 * k1.x + Da124 * k4.x + Da125 * k5.x + Da126 * k6.x + Da127 * k7.x + Da128 * k8.x + Da129 * k9.x + Da1210 * k10.x + Da1211 * k2.x),
                y.y + h * (Da121 * k1.y + Da124 * k4.y + Da125 * k5.y + Da126 * k6.y + Da127 * k7.y + Da128 * k8.y + Da129 * k9.y + Da1210 * k10.y + Da1211 * k2.y),
                y.z + h * (Da121 * k1.z + Da124 * k4.z + Da125 * k5.z + Da126 * k6.z + Da127 * k7.z + Da128 * k8.z + Da129 * k9.z + Da1210 * k10.z + Da1211 * k2.z));
        k3 = BLOCH(xph, yy1);

        ST(k4, Db1 * k1.x + Db6 * k6.x + Db7 * k7.x + Db8 * k8.x + Db9 * k9.x + Db10 * k10.x + Db11 * k2.x + Db12 * k3.x,
               Db1 * k1.y + Db6 * k6.y + Db7 * k7.y + Db8 * k8.y + Db9 * k9.y + Db10 * k10.y + Db11 * k2.y + Db12 * k3.y,
               Db1 * k1.z + Db6 * k6.z + Db7 * k7.z + Db8 * k8.z + Db9 * k9.z + Db10 * k10.z + Db11 * k2.z + Db12 * k3.z);
        ST(k5, y.x + h * k4.x, y.y + h * k4.y, y.z + h * k4.z);

        /* error estimation (:191-205) */
        or_coords sk, erri, sq;
        ST(sk, opt->atol + opt->rtol * max_d(fabs(y.x), fabs(k5.x)),
               opt->atol + opt->rtol * max_d(fabs(y.y), fabs(k5.y)),
               opt->atol + opt->rtol * max_d(fabs(y.z), fabs(k5.z)));
        ST(erri, k4.x - Dbhh1 * k1.x - Dbhh2 * k9.x - Dbhh3 * k3.x,
                 k4.y - Dbhh1 * k1.y - Dbhh2 * k9.y - Dbhh3 * k3.y,
                 k4.z - Dbhh1 * k1.z - Dbhh2 * k9.z - Dbhh3 * k3.z);
        ST(sq, erri.x / sk.x, erri.y / sk.y, erri.z / sk.z);
        err2 = 0.0 + (sq.x * sq.x + sq.y * sq.y + sq.z * sq.z);
        ST(erri, Der1 * k1.x + Der6 * k6.x + Der7 * k7.x + Der8 * k8.x + Der9 * k9.x + Der10 * k10.x + Der11 * k2.x + Der12 * k3.x,
                 Der1 * k1.y + Der6 * k6.y + Der7 * k7.y + Der8 * k8.y + Der9 * k9.y + Der10 * k10.y + Der11 * k2.y + Der12 * k3.y,
                 Der1 * k1.z + Der6 * k6.z + Der7 * k7.z + Der8 * k8.z + Der9 * k9.z + Der10 * k10.z + Der11 * k2.z + Der12 * k3.z);
        ST(sq, erri.x / sk.x, erri.y / sk.y, erri.z / sk.z);
        err = 0.0 + (sq.x * sq.x + sq.y * sq.y + sq.z * sq.z);
        deno = err + 0.01 * err2;
        if (deno <= 0.0) deno = 1.0;
        err = fabs(h) * err * sqrt(1.0 / (deno * (double)n));

        /* step-size control (:208-213) */
        fac11 = pow(err, expo1);
        fac = fac11 / pow(facold, opt->beta);
        fac = max_d(facc2, min_d(facc1, fac / opt->safe));
        hnew = h / fac;
        if (err <= 1.0 || opt->fixedStepSize) {
            facold = max_d(err, 1.0E-4);
            if (n_acc) (*n_acc)++;
            k4 = BLOCH(xph, k5);
            k1 = k4;
            y = k5;
            x = xph;
            if (last) { ret = 1; break; }
            if (fabs(hnew) > opt->hmax) hnew = posneg * opt->hmax;
            if (reject) hnew = posneg * min_d(fabs(hnew), fabs(h));
            reject = 0;
        } else {
            if (n_rej) (*n_rej)++;
            hnew = h / min_d(facc1, fac11 / opt->safe);
            reject = 1;
            last = 0;
        }
        if (!opt->fixedStepSize) h = hnew;
#undef C2T
#undef ST
    }
    *yio = y;
    *hio = h;
    return ret;
}

/* ------------------------------------------------------------------ */
/* RK45: src/integrator.cpp:251-337, tableau include/coeff.h:178-218    */
/* ------------------------------------------------------------------ */

int or_integrate_rk45(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej) {
    const double A2 = 2.0 / 9.0, A3 = 1.0 / 3.0, A4 = 3.0 / 4.0, A5 = 1.0, A6 = 5.0 / 6.0;
    const double B21 = 2.0 / 9.0;
    const double B31 = 1.0 / 12.0, B32 = 1.0 / 4.0;
    const double B41 = 69.0 / 128.0, B42 = -243.0 / 128.0, B43 = 135.0 / 64.0;
    const double B51 = -17.0 / 12.0, B52 = 27.0 / 4.0, B53 = -27.0 / 5.0, B54 = 16.0 / 15.0;
    const double B61 = 65.0 / 432.0, B62 = -5.0 / 16.0, B63 = 13.0 / 16.0, B64 = 4.0 / 27.0,
                 B65 = 5.0 / 144.0;
    const double CH1 = 47.0 / 450.0, CH2 = 0.0, CH3 = 12.0 / 25.0, CH4 = 32.0 / 225.0,
                 CH5 = 1.0 / 30.0, CH6 = 6.0 / 25.0;
    const double CT1 = 1.0 / 150.0, CT2 = 0.0, CT3 = -3.0 / 100.0, CT4 = 16.0 / 75.0,
                 CT5 = 1.0 / 20.0, CT6 = -6.0 / 25.0;

    or_coords y = *yio;
    double h = *hio;
    double t = t0;
    bool stop = false;
    int nstep = 0;
    double endOfSimulDt;
    double lastH = h;
    /* coords members default to zero (include/double3.h:17-21): the first
     * trial's k1 is evaluated at this zero vector (:298). */
    or_coords k1, k2, k3, k4, k5, k6, yy1 = {0.0, 0.0, 0.0}, TE2, ws;
    double error, tol, ratio, q;
    const double beta1 = 0.7, beta2 = -0.4, accept_safety = 0.81, kk = 7.0;
    double prev_ratio = 1.0;
    int ret = 0;
    for (;;) {
        nstep++;
        if ((unsigned int)nstep > opt->nmax) { ret = -2; break; }
        endOfSimulDt = tf - t;
        lastH = h;
        if (opt->fixedStepSize) {
            h = opt->h;
        } else {
            if (h > opt->hmax) {
                h = std_min(h, opt->hmax);
            } else if (h < opt->hmin) {
                ret = -1;
                break;
            }
        }
        if (h >= endOfSimulDt) {
            h = endOfSimulDt;
            stop = true;
        } else {
            lastH = h;
            stop = false;
        }
        k1 = BLOCH(t, yy1);
        { const double c = h * B21;
          yy1.x = y.x + c * k1.x; yy1.y = y.y + c * k1.y; yy1.z = y.z + c * k1.z; }
        k2 = BLOCH(t + A2 * h, yy1);
        { const double c1 = h * B31, c2 = h * B32;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z; }
        k3 = BLOCH(t + A3 * h, yy1);
        { const double c1 = h * B41, c2 = h * B42, c3 = h * B43;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x + c3 * k3.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y + c3 * k3.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z + c3 * k3.z; }
        k4 = BLOCH(t + A4 * h, yy1);
        { const double c1 = h * B51, c2 = h * B52, c3 = h * B53, c4 = h * B54;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x + c3 * k3.x + c4 * k4.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y + c3 * k3.y + c4 * k4.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z + c3 * k3.z + c4 * k4.z; }
        k5 = BLOCH(t + A5 * h, yy1);
        { const double c1 = h * B61, c2 = h * B62, c3 = h * B63, c4 = h * B64, c5 = h * B65;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x + c3 * k3.x + c4 * k4.x + c5 * k5.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y + c3 * k3.y + c4 * k4.y + c5 * k5.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z + c3 * k3.z + c4 * k4.z + c5 * k5.z; }
        k6 = BLOCH(t + A6 * h, yy1);
        ws.x = y.x + h * (k1.x * CH1 + k2.x * CH2 + k3.x * CH3 + k4.x * CH4 + k5.x * CH5 + k6.x * CH6);
        ws.y = y.y + h * (k1.y * CH1 + k2.y * CH2 + k3.y * CH3 + k4.y * CH4 + k5.y * CH5 + k6.y * CH6);
        ws.z = y.z + h * (k1.z * CH1 + k2.z * CH2 + k3.z * CH3 + k4.z * CH4 + k5.z * CH5 + k6.z * CH6);
        TE2.x = h * (CT1 * k1.x + CT2 * k2.x + CT3 * k3.x + CT4 * k4.x + CT5 * k5.x + CT6 * k6.x);
        TE2.y = h * (CT1 * k1.y + CT2 * k2.y + CT3 * k3.y + CT4 * k4.y + CT5 * k5.y + CT6 * k6.y);
        TE2.z = h * (CT1 * k1.z + CT2 * k2.z + CT3 * k3.z + CT4 * k4.z + CT5 * k5.z + CT6 * k6.z);
        error = len3(TE2);
        error = std_max(error, 1.0E-16);
        tol = opt->rtol;
        ratio = tol / error;
        q = pow(ratio, beta1 / kk) * pow(prev_ratio, beta2 / kk);
        q = std_min(q, 4.0);
        if ((q > accept_safety && error < 2 * tol) || opt->fixedStepSize) {
            if (n_acc) (*n_acc)++;
            y = ws;
            t += h;
            if (stop) {
                h = lastH;
                ret = 0;
                break;
            }
        } else {
            if (n_rej) (*n_rej)++;
            if (stop) stop = false;
        }
        if (!opt->fixedStepSize) h = h * q;
        prev_ratio = ratio;
    }
    *yio = y;
    *hio = h;
    return ret;
}

/* ------------------------------------------------------------------ */
/* The other adaptive integrators: integrateRKF45 (src/integrator.cpp: */
/* 339-426), integrateRK45Quaternion (:428-522), integrateRKF45-      */
/* Quaternion (:524-618) and integrateMagnusCFET (:620-723).          */
/* ------------------------------------------------------------------ */

/* The step-size bounds these four share (e.g. :355-385).  Unlike RK45's
 * (:267-292) the end-of-interval test comes first and skips the hmax/hmin
 * checks, and a fixed step size never sets `stop`: with fixedStepSize the
 * call runs until nstep exceeds nmax and returns -2, as the reference does.
 * Returns -1 (h below hmin) or 0. */
static int bounds_v2(const or_options *opt, double t, double tf, double *h, double *lastH,
        bool *stop) {
    const double endOfSimulDt = tf - t;
    *lastH = *h;
    if (opt->fixedStepSize) {
        *h = opt->h;
    } else if (*h >= endOfSimulDt) {
        *h = endOfSimulDt;
        *stop = true;
    } else {
        *lastH = *h;
        *stop = false;
        if (*h > opt->hmax) *h = std_min(*h, opt->hmax);
        else if (*h < opt->hmin) return -1;
    }
    return 0;
}

/* The controller and accept test shared by all of them (e.g. :300-326).
 * Returns 1 when the call ends (an accepted step with stop set). */
static int control(const or_options *opt, double error, double *prev_ratio, double *t,
        double *h, double lastH, bool *stop, or_coords *y, or_coords ynew,
        uint64_t *n_acc, uint64_t *n_rej) {
    error = std_max(error, 1.0E-16);
    const double tol = opt->rtol;
    const double ratio = tol / error;
    double q = pow(ratio, 0.7 / 7.0) * pow(*prev_ratio, -0.4 / 7.0);
    q = std_min(q, 4.0);
    if ((q > 0.81 && error < 2 * tol) || opt->fixedStepSize) {
        if (n_acc) (*n_acc)++;
        *y = ynew;
        *t += *h;
        if (*stop) {
            *h = lastH;
            return 1;
        }
    } else {
        if (n_rej) (*n_rej)++;
        if (*stop) *stop = false;
    }
    if (!opt->fixedStepSize) *h = *h * q;
    *prev_ratio = ratio;
    return 0;
}

/* integrateRKF45 (:339-426), Fehlberg tableau include/coeff.h:220-256.
 * As in RK45 the first stage is evaluated at the previous trial's sixth-stage
 * input (zero on the first trial), and the "error" is the length of the
 * fourth-order increment h*sum(CR_i k_i) (coefficients CR are the
 * fourth-order weights, :398), not a difference of two solutions. */
int or_integrate_rkf45(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej) {
    const double A2 = 1.0 / 4.0, A3 = 3.0 / 8.0, A4 = 12.0 / 13.0, A5 = 1.0, A6 = 0.5;
    const double B21 = 0.25;
    const double B31 = 3.0 / 32.0, B32 = 9.0 / 32.0;
    const double B41 = 1932.0 / 2197.0, B42 = -7200.0 / 2197.0, B43 = 7296.0 / 2197.0;
    const double B51 = 439.0 / 216.0, B52 = -8.0, B53 = 3680.0 / 513.0, B54 = -845.0 / 4104.0;
    const double B61 = -8.0 / 27.0, B62 = 2.0, B63 = -3544.0 / 2565.0, B64 = 1859.0 / 4104.0,
                 B65 = -11.0 / 40.0;
    const double C1 = 16.0 / 135.0, C2 = 0.0, C3 = 6656.0 / 12825.0, C4 = 28561.0 / 56430.0,
                 C5 = -9.0 / 50.0, C6 = 2.0 / 55.0;
    const double R1 = 25.0 / 216.0, R2 = 0.0, R3 = 1408.0 / 2565.0, R4 = 2197.0 / 4104.0,
                 R5 = -0.2, R6 = 0.0;
    or_coords y = *yio;
    double h = *hio, t = t0, lastH = h, prev_ratio = 1.0;
    bool stop = false;
    unsigned int nstep = 0;
    or_coords k1, k2, k3, k4, k5, k6, yy1 = {0.0, 0.0, 0.0}, TE2, ws;
    int ret = 0;
    for (;;) {
        nstep++;
        if (nstep > opt->nmax) { ret = -2; break; }
        if (bounds_v2(opt, t, tf, &h, &lastH, &stop)) { ret = -1; break; }
        k1 = BLOCH(t, yy1);
        { const double c = h * B21;
          yy1.x = y.x + c * k1.x; yy1.y = y.y + c * k1.y; yy1.z = y.z + c * k1.z; }
        k2 = BLOCH(t + A2 * h, yy1);
        { const double c1 = h * B31, c2 = h * B32;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z; }
        k3 = BLOCH(t + A3 * h, yy1);
        { const double c1 = h * B41, c2 = h * B42, c3 = h * B43;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x + c3 * k3.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y + c3 * k3.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z + c3 * k3.z; }
        k4 = BLOCH(t + A4 * h, yy1);
        { const double c1 = h * B51, c2 = h * B52, c3 = h * B53, c4 = h * B54;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x + c3 * k3.x + c4 * k4.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y + c3 * k3.y + c4 * k4.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z + c3 * k3.z + c4 * k4.z; }
        k5 = BLOCH(t + A5 * h, yy1);
        { const double c1 = h * B61, c2 = h * B62, c3 = h * B63, c4 = h * B64, c5 = h * B65;
          yy1.x = y.x + c1 * k1.x + c2 * k2.x + c3 * k3.x + c4 * k4.x + c5 * k5.x;
          yy1.y = y.y + c1 * k1.y + c2 * k2.y + c3 * k3.y + c4 * k4.y + c5 * k5.y;
          yy1.z = y.z + c1 * k1.z + c2 * k2.z + c3 * k3.z + c4 * k4.z + c5 * k5.z; }
        k6 = BLOCH(t + A6 * h, yy1);
        ws.x = y.x + h * (k1.x * C1 + k2.x * C2 + k3.x * C3 + k4.x * C4 + k5.x * C5 + k6.x * C6);
        ws.y = y.y + h * (k1.y * C1 + k2.y * C2 + k3.y * C3 + k4.y * C4 + k5.y * C5 + k6.y * C6);
        ws.z = y.z + h * (k1.z * C1 + k2.z * C2 + k3.z * C3 + k4.z * C4 + k5.z * C5 + k6.z * C6);
        TE2.x = h * (R1 * k1.x + R2 * k2.x + R3 * k3.x + R4 * k4.x + R5 * k5.x + R6 * k6.x);
        TE2.y = h * (R1 * k1.y + R2 * k2.y + R3 * k3.y + R4 * k4.y + R5 * k5.y + R6 * k6.y);
        TE2.z = h * (R1 * k1.z + R2 * k2.z + R3 * k3.z + R4 * k4.z + R5 * k5.z + R6 * k6.z);
        if (control(opt, len3(TE2), &prev_ratio, &t, &h, lastH, &stop, &y, ws, n_acc, n_rej))
            break;
    }
    *yio = y;
    *hio = h;
    return ret;
}

/* Quaternions, src/quaternion.cpp:128-164 */
typedef struct { double w, x, y, z; } or_quat;

static or_quat q_mult(or_quat a, or_quat b) {
    or_quat o;
    o.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    o.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    o.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    o.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return o;
}

/* q v q* (:137-148) */
static or_coords qv_mult(or_quat q, or_coords v) {
    const or_quat p = {0.0, v.x, v.y, v.z};
    const or_quat c = {q.w, -q.x, -q.y, -q.z};
    const or_quat r = q_mult(q_mult(q, p), c);
    const or_coords o = {r.x, r.y, r.z};
    return o;
}

/* The rotation by |k| dt about -k/|k| (:150-156) */
static or_quat rodriguez_quat(or_coords k, double dt) {
    const double angle = len3(k);
    const or_coords n = {k.x / angle, k.y / angle, k.z / angle};
    const double hh = angle * dt;
    const double s = -sin(hh / 2.0);
    const or_quat o = {cos(hh / 2.0), n.x * s, n.y * s, n.z * s};
    return o;
}

/* v rotated by |k| about -k/|k| (:158-163) */
static or_coords rodriguez(or_coords k, or_coords v) {
    const double angle = len3(k);
    const double s = sin(angle);
    const double c = cos(angle);
    const or_coords vk = cross3(v, k);
    const double sa = s / angle;
    const double f = (dot3(k, v) * (1.0 - c)) / (angle * angle);
    const or_coords o = {(v.x * c + vk.x * sa) + k.x * f, (v.y * c + vk.y * sa) + k.y * f,
                         (v.z * c + vk.z * sa) + k.z * f};
    return o;
}

/* integrateRK45Quaternion / integrateRKF45Quaternion: six field samples
 * k_i = findCrossTerm(t + A_i h); the step is y rotated by the product of
 * rotations R(k_i, h W_i), i = 1..6, applied k1 first; the comparison
 * solution drops the sixth factor and uses weights E_i.  RK45 tableau:
 * A from RK45COEF, W = CH, E = C; RKF45 tableau: A from RKF45COEF, W = E = C
 * (include/coeff.h:178-256). */
static int integrate_quat(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej, const double *A, const double *W, const double *E) {
    or_coords y = *yio;
    double h = *hio, t = t0, lastH = h, prev_ratio = 1.0;
    bool stop = false;
    unsigned int nstep = 0;
    int ret = 0;
    for (;;) {
        nstep++;
        if (nstep > opt->nmax) { ret = -2; break; }
        if (bounds_v2(opt, t, tf, &h, &lastH, &stop)) { ret = -1; break; }
        or_coords k[6];
        k[0] = or_cross_term(t, opt, ext, t0, p_old, v_old);
        for (int i = 1; i < 6; i++) k[i] = or_cross_term(t + A[i] * h, opt, ext, t0, p_old, v_old);
        or_quat qw = rodriguez_quat(k[0], h * W[0]);
        or_quat qe = rodriguez_quat(k[0], h * E[0]);
        for (int i = 1; i < 5; i++) {
            qw = q_mult(rodriguez_quat(k[i], h * W[i]), qw);
            qe = q_mult(rodriguez_quat(k[i], h * E[i]), qe);
        }
        qw = q_mult(rodriguez_quat(k[5], h * W[5]), qw);
        const or_coords ws = qv_mult(qw, y);
        const or_coords te = qv_mult(qe, y);
        const or_coords d = {ws.x - te.x, ws.y - te.y, ws.z - te.z};
        if (control(opt, len3(d), &prev_ratio, &t, &h, lastH, &stop, &y, ws, n_acc, n_rej))
            break;
    }
    *yio = y;
    *hio = h;
    return ret;
}

int or_integrate_rk45q(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej) {
    static const double A[6] = {0.0, 2.0 / 9.0, 1.0 / 3.0, 3.0 / 4.0, 1.0, 5.0 / 6.0};
    static const double W[6] = {47.0 / 450.0, 0.0, 12.0 / 25.0, 32.0 / 225.0, 1.0 / 30.0, 6.0 / 25.0};
    static const double E[5] = {1.0 / 9.0, 0.0, 9.0 / 20.0, 16.0 / 45.0, 1.0 / 12.0};
    return integrate_quat(t0, tf, yio, p_old, v_old, opt, ext, hio, n_acc, n_rej, A, W, E);
}

int or_integrate_rkf45q(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej) {
    static const double A[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 0.5};
    static const double W[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0,
                                -9.0 / 50.0, 2.0 / 55.0};
    return integrate_quat(t0, tf, yio, p_old, v_old, opt, ext, hio, n_acc, n_rej, A, W, W);
}

/* integrateMagnusCFET (:620-723): the commutator-free exponential
 * integrator of order 8 (11 exponentials) with an embedded order-6 one (5
 * exponentials) from Alvermann & Fehske, J. Comput. Phys. 230 (2011) 5930
 * (arXiv:1102.5071), on the field sampled at the 5 Gauss-Legendre nodes;
 * coefficients include/coeff.h:258-321 (quadrature weights folded in). */
static const double GL5X[5] = {0.046910077030668074, 0.2307653449471585, 0.5,
                               0.7692346550528415, 0.9530899229693319};
static const double CF8[6][5] = {
    {-0.005984812750244602, 0.020217897716672675, -0.03646660716686794, 0.04245408899634467,
     0.14949496424802844},
    {0.02512881778391826, -0.08955337580937157, 0.19032828600393248, 0.28920194841768715,
     -0.03568486888016069},
    {-0.04972613261435144, 0.6572496090804754, 0.11725149008224456, -0.43611911073577403,
     0.18080345083145516},
    {0.0464837870954061, -0.6265183532355708, -0.11144873822286155, 0.4154988625423615,
     -0.17224148557040517},
    {0.015567390500242426, -0.0677120111619075, -0.15241336773272682, -0.08803798509100481,
     -0.0013284996209207986},
    {-0.034049168494873015, 0.12263276452977083, 0.2699423185170027, 0.12263276452977094,
     -0.03404916849487304}};
static const double CF6[3][5] = {
    {-0.007431519425246311, 0.022452824849980513, -0.03620393417193995, 0.0409779901800809,
     0.1516046385671249},
    {0.03095749686312404, -0.10116885619385878, 0.1997455955716367, 0.3010242365525433,
     -0.055594729593982636},
    {-0.001072443882925378, -0.023971860139062367, -0.04263887835494916, -0.023971860139062388,
     -0.0010724438829253797}};

/* (c_1 B1 + c_2 B2 + c_3 B3 + c_4 B4 + c_5 B5) * h, summed left to right;
 * rev takes the row backwards (c_5 B1 + ... + c_1 B5) */
static or_coords cfet_exp(const double *c, int rev, const or_coords *B, double h) {
    double cc[5];
    for (int i = 0; i < 5; i++) cc[i] = rev ? c[4 - i] : c[i];
    or_coords o;
    o.x = ((((cc[0] * B[0].x + cc[1] * B[1].x) + cc[2] * B[2].x) + cc[3] * B[3].x) + cc[4] * B[4].x) * h;
    o.y = ((((cc[0] * B[0].y + cc[1] * B[1].y) + cc[2] * B[2].y) + cc[3] * B[3].y) + cc[4] * B[4].y) * h;
    o.z = ((((cc[0] * B[0].z + cc[1] * B[1].z) + cc[2] * B[2].z) + cc[3] * B[3].z) + cc[4] * B[4].z) * h;
    return o;
}

int or_integrate_magnus(double t0, double tf, or_coords *yio, or_coords p_old,
        or_coords v_old, const or_options *opt, const or_ext *ext, double *hio,
        uint64_t *n_acc, uint64_t *n_rej) {
    or_coords y = *yio;
    double h = *hio, t = t0, lastH = h, prev_ratio = 1.0;
    bool stop = false;
    unsigned int nstep = 0;
    int ret = 0;
    for (;;) {
        nstep++;
        if (nstep > opt->nmax) { ret = -2; break; }
        if (bounds_v2(opt, t, tf, &h, &lastH, &stop)) { ret = -1; break; }
        or_coords B[5];
        for (int i = 0; i < 5; i++) B[i] = or_cross_term(t + GL5X[i] * h, opt, ext, t0, p_old, v_old);
        /* order 8: k11 = row 1 reversed is applied first, then rows 2..5
         * reversed (k10..k7), row 6 (k6), rows 5..1 forward (k5..k1) */
        or_coords y8 = y;
        for (int r = 0; r < 5; r++) y8 = rodriguez(cfet_exp(CF8[r], 1, B, h), y8);
        y8 = rodriguez(cfet_exp(CF8[5], 0, B, h), y8);
        for (int r = 4; r >= 0; r--) y8 = rodriguez(cfet_exp(CF8[r], 0, B, h), y8);
        /* order 6: k5 = row 1 reversed, k4 = row 2 reversed, k3 = row 3, k2 =
         * row 2, k1 = row 1 */
        or_coords y6 = y;
        y6 = rodriguez(cfet_exp(CF6[0], 1, B, h), y6);
        y6 = rodriguez(cfet_exp(CF6[1], 1, B, h), y6);
        y6 = rodriguez(cfet_exp(CF6[2], 0, B, h), y6);
        y6 = rodriguez(cfet_exp(CF6[1], 0, B, h), y6);
        y6 = rodriguez(cfet_exp(CF6[0], 0, B, h), y6);
        const or_coords d = {y8.x - y6.x, y8.y - y6.y, y8.z - y6.z};
        if (control(opt, len3(d), &prev_ratio, &t, &h, lastH, &stop, &y, y8, n_acc, n_rej))
            break;
    }
    *yio = y;
    *hio = h;
    return ret;
}
#undef BLOCH

/* ------------------------------------------------------------------ */
/* Spectral path (integratorType 6)                                     */
/* ------------------------------------------------------------------ */

/* findNoiseTerm, src/integrator.cpp:72-80 (+ testNoise :36-38):
 * gamma * ((1/c^2) v x E + G.r + a * sin(w t)), no pulse and no B0 */
or_coords or_noise_term(double t, const or_options *opt, double t0, or_coords p_old,
        or_coords v_old) {
    or_coords p, v;
    or_interpolate(t, t0, p_old, v_old, &p, &v, opt);
    const or_coords G = field_grad(p, opt);
    or_coords N;
    N.x = opt->noiseAmplitudes.x * sin(opt->noiseFrequencies.x * t);
    N.y = opt->noiseAmplitudes.y * sin(opt->noiseFrequencies.y * t);
    N.z = opt->noiseAmplitudes.z * sin(opt->noiseFrequencies.z * t);
    const or_coords vxE = cross3(v, opt->E);
    const double ic2 = 1.0 / C2;
    or_coords w;
    w.x = ((vxE.x * ic2 + G.x) + N.x) * opt->gamma;
    w.y = ((vxE.y * ic2 + G.y) + N.y) * opt->gamma;
    w.z = ((vxE.z * ic2 + G.z) + N.z) * opt->gamma;
    return w;
}

/* SpectrumAggregator::update, src/floquet.cpp:111-118, with
 * goertzel_stage_1 (:25-31), goertzel_stage_2_vector (:33-38) and
 * covMat::add_outer(a, b, x, 0) (:105-108), which only touches imag_diag */
void or_spectrum_update(or_specagg *a, const or_speccfg *c, or_coords x) {
    for (int i = 0; i < OR_NW; i++) {
        const double angle = c->w[i] * c->dt;
        const double k = 2 * cos(angle);
        or_coords ns;
        ns.x = (x.x + k * a->s1[i].x) - a->s2[i].x;
        ns.y = (x.y + k * a->s1[i].y) - a->s2[i].y;
        ns.z = (x.z + k * a->s1[i].z) - a->s2[i].z;
        a->s2[i] = a->s1[i];
        a->s1[i] = ns;
        const double cw = cos(angle), sw = sin(angle);
        or_coords re, im;
        re.x = a->s1[i].x - cw * a->s2[i].x;
        re.y = a->s1[i].y - cw * a->s2[i].y;
        re.z = a->s1[i].z - cw * a->s2[i].z;
        im.x = sw * a->s2[i].x;
        im.y = sw * a->s2[i].y;
        im.z = sw * a->s2[i].z;
        a->im[i].x = a->im[i].x + 2 * (im.x * x.x - re.x * 0.0);
        a->im[i].y = a->im[i].y + 2 * (im.y * x.y - re.y * 0.0);
        a->im[i].z = a->im[i].z + 2 * (im.z * x.z - re.z * 0.0);
    }
    a->n_samples += 1;
}

/* first_sample_point, src/integrator.cpp:724-728 */
static double first_sample_point(double t0, double h) { return (floor(t0 / h) + 1) * h; }

/* integrateSpectrum, src/integrator.cpp:730-743: the noise field at the
 * multiples of h inside (t0, tf], halved, into the aggregator */
int or_integrate_spectrum(double t0, double tf, or_specagg *agg, const or_speccfg *cfg,
        or_coords p_old, or_coords v_old, const or_options *opt, double h) {
    double t = first_sample_point(t0, h);
    const double next_t = first_sample_point(tf, h);
    int n_steps = 0;
    while (t < next_t - (h / 2)) {
        const or_coords B = or_noise_term(t, opt, t0, p_old, v_old);
        const or_coords x = {B.x / 2, B.y / 2, B.z / 2};
        or_spectrum_update(agg, cfg, x);
        t += h;
        n_steps++;
    }
    return n_steps;
}

/* SpectrumAggregator::get_covariance_spectrum, src/floquet.cpp:131-147
 * (LAPLACE): per frequency the 3x3 matrix c c^H, c = a + i b from
 * goertzel_stage_2, plus -i imag_diag on the diagonal; out[i][r][s] = (re, im) */
void or_covariance(const or_specagg *a, const or_speccfg *cfg, double out[OR_NW][3][3][2]) {
    for (int i = 0; i < OR_NW; i++) {
        const double angle = cfg->w[i] * cfg->dt;
        const double cw = cos(angle), sw = sin(angle);
        const double s1[3] = {a->s1[i].x, a->s1[i].y, a->s1[i].z};
        const double s2[3] = {a->s2[i].x, a->s2[i].y, a->s2[i].z};
        const double id[3] = {a->im[i].x, a->im[i].y, a->im[i].z};
        double re[3], im[3];
        for (int r = 0; r < 3; r++) {
            re[r] = s1[r] - cw * s2[r];
            im[r] = sw * s2[r];
        }
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) {
                /* (re_r + i im_r) * conj(re_q + i im_q) */
                out[i][r][q][0] = re[r] * re[q] - im[r] * (-im[q]);
                out[i][r][q][1] = re[r] * (-im[q]) + im[r] * re[q];
            }
        for (int r = 0; r < 3; r++) {
            out[i][r][r][0] = out[i][r][r][0] + 0.0;
            out[i][r][r][1] = out[i][r][r][1] + (-id[r]);
        }
    }
}

/* integrateHamiltonian, src/integrator.cpp:745-789: the order-8 CFET
 * propagator (11 exponentials of the field at the Gauss-Legendre nodes) in
 * fixed steps h (the last one shortened to end at tf), along the field at
 * the origin at rest (p_old = v_old = 0); y <- q1 q2 ... q11 y */
int or_integrate_hamiltonian(double t0, double tf, double *qio, const or_options *opt,
        const or_ext *ext, double h) {
    or_quat y = {qio[0], qio[1], qio[2], qio[3]};
    const or_coords zero = {0.0, 0.0, 0.0};
    double t = t0;
    unsigned int nstep = 0;
    int ret = 0;
    while (t < tf) {
        nstep++;
        if (nstep > opt->nmax) { ret = -2; break; }
        const double rest = tf - t;
        if (h >= rest) h = rest;
        or_coords B[5];
        for (int i = 0; i < 5; i++) B[i] = or_cross_term(t + GL5X[i] * h, opt, ext, t0, zero, zero);
        or_quat q[11];   /* q[k] = q_{k+1} */
        for (int r = 0; r < 6; r++) q[r] = rodriguez_quat(cfet_exp(CF8[r], 0, B, 1.0), h);
        for (int r = 4; r >= 0; r--) q[10 - r] = rodriguez_quat(cfet_exp(CF8[r], 1, B, 1.0), h);
        or_quat P = q[0];
        for (int k = 1; k < 11; k++) P = q_mult(P, q[k]);
        y = q_mult(P, y);
        t += h;
    }
    qio[0] = y.w; qio[1] = y.x; qio[2] = y.y; qio[3] = y.z;
    return ret;
}

/* ------------------------------------------------------------------ */
/* Ensemble: src/particle.cpp:638-837                                   */
/* ------------------------------------------------------------------ */

/* src/particle.cpp:638-710 (CPU variant).  partID and wall_hit, which the
 * reference leaves uninitialised, are set to the global index and 0. */
void or_init_particles(const or_options *opt, const or_arrays *a, int n,
        unsigned int first, unsigned int stride) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int j = 0; j < n; j++) {
        const unsigned int ipart = first + (unsigned int)j * stride;
        or_rng rng;
        or_seed(&rng, opt->seed, ipart);
        or_coords p;
        p.x = or_uniform(&rng) * opt->L.x - opt->L.x / 2.0;
        p.y = or_uniform(&rng) * opt->L.y - opt->L.y / 2.0;
        p.z = or_uniform(&rng) * opt->L.z - opt->L.z / 2.0;
        a->pos[j] = p;
        a->pos_old[j] = p;
        if (opt->gas_coll)
            a->next_gas[j] = or_exponential(&rng, opt->tc);
        else
            a->next_gas[j] = DBL_MAX;
        or_coords vv = {0.0, 0.0, 0.0};
        if (opt->dist == 'C') {
            or_coords vec;
            vec.x = or_normal(&rng, 0.0, 1.0);
            vec.y = or_normal(&rng, 0.0, 1.0);
            vec.z = or_normal(&rng, 0.0, 1.0);
            double vn = len3(vec);
            vv.x = (opt->V * vec.x) / vn;
            vv.y = (opt->V * vec.y) / vn;
            vv.z = (opt->V * vec.z) / vn;
        } else if (opt->dist == 'M') {
            vv.x = or_maxboltz(&rng, opt->sqrtKT_m);
            vv.y = or_maxboltz(&rng, opt->sqrtKT_m);
            vv.z = or_maxboltz(&rng, opt->sqrtKT_m);
        }
        a->v[j] = vv;
        a->v_old[j] = vv;
        a->state[j] = rng;
        a->S[j] = opt->yi;
        a->t[j] = opt->t0;
        a->t_old[j] = opt->t0;
        a->tf[j] = opt->tf;
        a->dt[j] = opt->h;
        a->h[j] = opt->h;
        a->stop[j] = false;
        a->coll_type[j] = 'W';
        a->wall_hit[j] = 0;
        a->failure[j] = 0;
        a->n_bounce[j] = 0;
        a->n_coll[j] = 0;
        a->n_steps[j] = 0;
        a->partID[j] = ipart;
    }
}

/* src/particle.cpp:712-837.  Types outside 0-6 leave the spin alone as the
 * reference's final else does.  Type 6 samples the noise field into agg[j]
 * (reset first, :746-747) and needs agg and cfg; it leaves S alone. */
static void run_simulation(const or_options *opt, const or_ext *ext, const or_arrays *a,
        int n, double nextTOut, const double *stop_at, double prevTOut,
        uint64_t *n_ode, int nthreads, or_specagg *agg, const or_speccfg *cfg) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#else
    (void)nthreads;
#endif
    for (int j = 0; j < n; j++) {
        if (agg) memset(agg + j, 0, sizeof *agg);
        double target = nextTOut;
        if (stop_at) {
            if (!(prevTOut < stop_at[j])) continue;
            if (stop_at[j] < target) target = stop_at[j];
        }
        or_coords S = a->S[j], v = a->v[j], v_old = a->v_old[j], pos = a->pos[j],
                  pos_old = a->pos_old[j];
        double t = a->t[j], t_old = a->t_old[j], tf, dt;
        double next_gas = a->next_gas[j];
        double h = a->h[j];
        or_rng state = a->state[j];
        uint64_t n_bounce = a->n_bounce[j], n_coll = a->n_coll[j], n_steps = a->n_steps[j];
        int failure = a->failure[j];
        bool stop = a->stop[j];
        char coll_type = a->coll_type[j], wall_hit = a->wall_hit[j];
        uint64_t nacc = 0, nrej = 0, nseg = 0;
        dt = target - t;
        tf = target;
        bool finished = false;
        int spinResult = 0;
        double tempH = h;
        or_specagg sa;
        if (agg) memset(&sa, 0, sizeof sa);
        while (!finished && !stop) {
            or_calc_next_collision_time(t, tf, v, pos, &next_gas, &dt, &coll_type,
                    &n_bounce, &n_coll, &finished, &wall_hit, &state, opt);
            or_update_position_and_velocity(&t_old, &t, &dt, &pos_old, &pos, &v_old, &v,
                    &coll_type, &wall_hit, &state, &stop, opt);
            or_sanity_check(&t_old, &t, &pos_old, &pos, &v, &v_old, &stop, &failure, opt);
            if (failure == 0) {
                if (opt->integratorType == 0)
                    spinResult = or_integrate_dop(t_old, t, &S, pos_old, v_old, opt, ext,
                            &tempH, &nacc, &nrej);
                else if (opt->integratorType == 1)
                    spinResult = or_integrate_rk45(t_old, t, &S, pos_old, v_old, opt, ext,
                            &tempH, &nacc, &nrej);
                else if (opt->integratorType == 2)
                    spinResult = or_integrate_magnus(t_old, t, &S, pos_old, v_old, opt, ext,
                            &tempH, &nacc, &nrej);
                else if (opt->integratorType == 3)
                    spinResult = or_integrate_rk45q(t_old, t, &S, pos_old, v_old, opt, ext,
                            &tempH, &nacc, &nrej);
                else if (opt->integratorType == 4)
                    spinResult = or_integrate_rkf45(t_old, t, &S, pos_old, v_old, opt, ext,
                            &tempH, &nacc, &nrej);
                else if (opt->integratorType == 5)
                    spinResult = or_integrate_rkf45q(t_old, t, &S, pos_old, v_old, opt, ext,
                            &tempH, &nacc, &nrej);
                else if (opt->integratorType == 6 && agg)
                    /* the CPU path drops the sample count (:777-778) */
                    (void)or_integrate_spectrum(t_old, t, &sa, cfg, pos_old, v_old, opt, tempH);
                else
                    spinResult = 0;
                nseg++;
                if (opt->keepStepSize) h = tempH;
                if (spinResult < 0) {
                    stop = true;
                    failure = spinResult;
                }
            }
            if (stop) {
                t = NAN;
                pos.x = NAN; pos.y = NAN; pos.z = NAN;
                v.x = NAN; v.y = NAN; v.z = NAN;
                S.x = NAN; S.y = NAN; S.z = NAN;
            }
            n_steps += 1;
        }
        a->S[j] = S;
        a->v[j] = v;
        a->v_old[j] = v_old;
        a->pos[j] = pos;
        a->pos_old[j] = pos_old;
        a->t[j] = t;
        a->t_old[j] = t_old;
        a->next_gas[j] = next_gas;
        a->h[j] = h;
        a->state[j] = state;
        a->failure[j] = failure;
        a->n_bounce[j] = n_bounce;
        a->n_coll[j] = n_coll;
        a->n_steps[j] = n_steps;
        a->stop[j] = stop;
        a->coll_type[j] = coll_type;
        a->wall_hit[j] = wall_hit;
        a->tf[j] = tf;
        if (agg) agg[j] = sa;
        if (n_ode) {
            n_ode[3 * (size_t)j + 0] += nacc;
            n_ode[3 * (size_t)j + 1] += nrej;
            n_ode[3 * (size_t)j + 2] += nseg;
        }
    }
}

void or_run_simulation(const or_options *opt, const or_ext *ext, const or_arrays *a,
        int n, double nextTOut, const double *stop_at, double prevTOut,
        uint64_t *n_ode, int nthreads) {
    run_simulation(opt, ext, a, n, nextTOut, stop_at, prevTOut, n_ode, nthreads, NULL, NULL);
}

void or_run_simulation_spec(const or_options *opt, const or_ext *ext, const or_arrays *a,
        int n, double nextTOut, const double *stop_at, double prevTOut,
        uint64_t *n_ode, int nthreads, or_specagg *agg, const or_speccfg *cfg) {
    run_simulation(opt, ext, a, n, nextTOut, stop_at, prevTOut, n_ode, nthreads, agg, cfg);
}

/* get_covariance_spectrum of particles 0..n-1 into out[n][OR_NW][3][3][2] */
void or_covariances(const or_specagg *agg, int n, const or_speccfg *cfg, double *out) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int j = 0; j < n; j++)
        or_covariance(agg + j, cfg, (double (*)[3][3][2])(out + (size_t)j * OR_NW * 18));
}

int or_sizeof_specagg(void) { return (int)sizeof(or_specagg); }
