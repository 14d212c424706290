// st_device.h -- per-lane device math of the spin-tracking hot path (gfx950).
//
// One particle per lane.  Everything here is FP64 on the vector ALUs.  The
// translation unit is compiled with -ffp-contract=off, so the "exact"
// functions round every +,-,*,/ exactly as the reference's x86-64 build
// does, in the reference's operand order (its coords operators are
// left-associative, src/double3.cpp:21-145).  The "fast" functions opt back
// into FMA contraction with a scoped pragma and restructure the right-hand
// side around a per-segment field polynomial; they agree with the exact
// path to the integrator tolerance, not bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spintrack_fields.h"

// Device math is also compiled for the host so a CPU build of the same
// functions can be checked against the oracle without a GPU.
#define ST_HD __host__ __device__

namespace st {

// ---------------------------------------------------------------------------
// constants (include/options.h:7-8)
// ---------------------------------------------------------------------------
constexpr double G_CONST = -9.81;
constexpr double HALF_G = 0.5 * G_CONST;                 // 0.5*a.y, exact
constexpr double C2 = 299792458.0 * 299792458.0;
constexpr double IC2 = 1.0 / C2;                          // the reference's 1.0/c2
constexpr double PI_D = 3.14159265358979323846;           // M_PI
constexpr double FLT_MIN_D = 1.17549435082228750797e-38;  // FLT_MIN
constexpr double DBL_MAX_D = 1.79769313486231570815e+308; // DBL_MAX

// arithmetic variants of the lane engine
enum { VAR_EXACT = 0, VAR_FAST = 1 };
// integrator types (options.integratorType; src/particle.cpp:568-592)
enum { INT_NONE = -1, INT_DOP853 = 0, INT_RK45 = 1, INT_MAGNUS = 2, INT_RK45Q = 3, INT_RKF45 = 4, INT_RKF45Q = 5,
       INT_SPECTRUM = 6 };

struct d3 {
    double x, y, z;
};

ST_HD __forceinline__ d3 mk3(double x, double y, double z) { return d3{x, y, z}; }

// src/double3.cpp:85-91
ST_HD __forceinline__ d3 cross3(const d3& a, const d3& b) {
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// src/double3.cpp:93-95
ST_HD __forceinline__ double dot3(const d3& a, const d3& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
// src/double3.cpp:101-103
ST_HD __forceinline__ double len3(const d3& a) {
    return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

// src/particle.cpp:24-31: sgn(0) = +1
ST_HD __forceinline__ double psgn(double v) { return (v < 0) ? -1.0 : 1.0; }
// std::min / std::max as called by the reference
ST_HD __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }
ST_HD __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }
// src/integrator.cpp:20-33
ST_HD __forceinline__ double sign_d(double a, double b) { return (b < 0.0) ? -fabs(a) : fabs(a); }
ST_HD __forceinline__ double min_d(double a, double b) { return (a < b) ? a : b; }
ST_HD __forceinline__ double max_d(double a, double b) { return (a > b) ? a : b; }

// ---------------------------------------------------------------------------
// Uniform (SGPR-resident) run parameters, derived once on the host from the
// reference options block.
// ---------------------------------------------------------------------------
// Physics of one field configuration / population: what the fields, the
// species and the wall/gas models are.  A plain handle has one (in the
// uniform kernel-argument block); a batched handle keeps one per
// configuration in HBM and every lane reads its particle's.
struct Phys {
    double B0x, B0y, B0z;
    double Ex, Ey, Ez;
    double G[9];          // rows Gx, Gy, Gz (grad, src/integrator.cpp:44-46)
    double gamma;
    double pa, pw;        // pulse amplitude and angular frequency
    double tc, sqrtKT_m, diffuse, V;
    double yix, yiy, yiz; // initial spin
    int32_t pulse_axis;
    int32_t pulse_on;     // 0: the pulse term is +0 and may be skipped bit-exactly
    int32_t diffuse_on;
    int32_t dist;         // 'C' or 'M'
    // The fast variant's segment polynomial (make_poly) in gamma-scaled form,
    // from the fields above and the gravity switch: gamma*G, gamma*B0,
    // gamma*E/c^2, the constant part gamma*(a x E)/c^2 of the linear term and
    // the quadratic term gamma*0.5*G.a, a = (0, -g, 0) or 0; and gamma*pa.
    double gG[9], gB0[3], gE[3], gw1[3], gw2[3], gpa;
};

// Run parameters shared by every particle of a launch (geometry, integrator
// control) plus the physics of a plain handle.
struct SimParams {
    double Lx, Ly, Lz;
    double rtol, atol, beta, safe, fac1, fac2, hmax, hmin, h0;
    double expo1, facc1, facc2;   // integrateDOP set-up (src/integrator.cpp:105-108)
    double maxPosStep;
    uint32_t nmax;
    int32_t pow_beta_on;  // 0: beta == 0, pow(facold, beta) == 1 exactly
    int32_t gravity, gas_coll, fixed, keep;
    Phys ph;
};

// ---------------------------------------------------------------------------
// RNG (src/particle.cpp:33-119)
// ---------------------------------------------------------------------------
struct Rng {
    uint64_t x, y, z, w;
    double spare;
    bool has;
};

// The reference's xoshiro variant (src/particle.cpp:59-70): output x+z, and
// w becomes rotl(z, 45) instead of rotl(w, 45).
ST_HD __forceinline__ uint64_t xo_next(Rng& s) {
    const uint64_t r = s.x + s.z;
    const uint64_t t = s.y << 17;
    s.z ^= s.x;
    s.w ^= s.y;
    s.y ^= s.z;
    s.x ^= s.w;
    s.z ^= t;
    s.w = (s.z << 45) | (s.z >> 19);
    return r;
}
// src/particle.cpp:72-83 (53-bit integer -> double is exact)
ST_HD __forceinline__ double uni(Rng& s) { return (double)(xo_next(s) >> 11) * 0x1.0p-53; }
// src/particle.cpp:85-87
ST_HD __forceinline__ double uni_lh(Rng& s, double lo, double hi) { return uni(s) * (hi - lo) + lo; }

// src/particle.cpp:89-107: Marsaglia polar method, spare cached in the state
ST_HD inline double normal(Rng& s, double mean, double sd) {
    if (s.has) {
        s.has = false;
        return s.spare * sd + mean;
    }
    double ts, tu, tv;
    do {
        tu = uni_lh(s, -1.0, 1.0);
        tv = uni_lh(s, -1.0, 1.0);
        ts = tu * tu + tv * tv;
    } while (ts >= 1.0 || ts == 0.0);
    ts = sqrt(-2.0 * log(ts) / ts);
    s.spare = tv * ts;
    s.has = true;
    return mean + sd * tu * ts;
}
ST_HD __forceinline__ double maxboltz(Rng& s, double sqrtkT_m) { return normal(s, 0.0, 1.0) * sqrtkT_m; }
// One polar-method pair, as normal() makes it when no spare is cached: the
// returned draw 0 + (1*tu)*ts (exactly tu*ts) and the new spare tv*ts.
ST_HD __forceinline__ void polar_pair(Rng& s, double& first, double& second) {
    double ts, tu, tv;
    do {
        tu = uni_lh(s, -1.0, 1.0);
        tv = uni_lh(s, -1.0, 1.0);
        ts = tu * tu + tv * tv;
    } while (ts >= 1.0 || ts == 0.0);
    ts = sqrt(-2.0 * log(ts) / ts);
    first = 0.0 + 1.0 * tu * ts;
    second = tv * ts;
}
// Three consecutive normal(s, 0, 1) draws.  With a spare cached they are
// spare, then one new pair (first, second); without, one pair and the first
// value of a second pair whose other value stays cached.  So a lane needs
// one or two pairs, always in the same order from its stream, and a wave runs
// the polar-method code twice at most (three calls of normal() would run it
// three times whenever the wave's lanes differ in their cached spare).  A
// spare is returned as spare*1 + 0, i.e. exactly the spare.
ST_HD __forceinline__ void normal3(Rng& s, double& n0, double& n1, double& n2) {
    const bool had = s.has;
    const double sp = s.spare;
    double a0, a1;
    polar_pair(s, a0, a1);
    if (had) {
        n0 = sp * 1.0 + 0.0;
        n1 = a0;
        n2 = a1 * 1.0 + 0.0;
        s.spare = a1;
        s.has = false;
    } else {
        double b0, b1;
        polar_pair(s, b0, b1);
        n0 = a0;
        n1 = a1 * 1.0 + 0.0;
        n2 = b0;
        s.spare = b1;
        s.has = true;
    }
}
ST_HD __forceinline__ double unif02pi(Rng& s) { return uni(s) * 2.0 * PI_D; }
ST_HD __forceinline__ double expo_draw(Rng& s, double tc) { return -tc * log(1.0 - uni(s)); }

// src/particle.cpp:646-663 / 455-473: splitmix64(seed + index), the index
// sum taken in 32-bit unsigned arithmetic, split into 32-bit halves, then
// three discarded draws.
ST_HD inline void seed_rng(Rng& s, uint32_t seed, uint32_t ipart) {
    uint64_t st = (uint64_t)(uint32_t)(seed + ipart);
    uint64_t r = (st += 0x9E3779B97f4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    r = r ^ (r >> 31);
    s.x = (uint32_t)r;
    s.y = (uint32_t)(r >> 32);
    r = (st += 0x9E3779B97f4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    r = r ^ (r >> 31);
    s.z = (uint32_t)r;
    s.w = (uint32_t)(r >> 32);
    s.spare = 0.0;
    s.has = false;
    (void)uni(s);
    (void)uni(s);
    (void)uni(s);
}

// ---------------------------------------------------------------------------
// Kinematics (src/particle.cpp:121-444)
// ---------------------------------------------------------------------------
struct Kin {
    double t, t_old, next_gas, dt;
    d3 pos, pos_old, v, v_old;
    uint64_t n_bounce, n_coll;
    int32_t failure;
    bool stop, finished;
    char coll_type, wall_hit;
};

// src/particle.cpp:121-257 -- closed-form wall hit times (quadratic in y
// under gravity), the else-if fix-up chain, tie rules, gas vs wall, and the
// maxPosStep / interval-end cap.  A gas collision's draw of the next gas
// time (:228) is left to the caller (gas_draw = true): the next draw from the
// particle's stream is the re-emission's, so drawing the gas time just before
// it keeps the stream order, and the two draws of a gas event sit together in
// gas_job below.  next_collision() is the whole function.
ST_HD inline void next_collision_core(const SimParams& P, const Phys& F, double tf, Kin& K, Rng& rng,
                                      bool& gas_draw) {
    const d3 v = K.v, pos = K.pos;
    const double t = K.t;
    double dx, dy, dz, dtx, dty, dtz;
    if (P.gravity) {
        dx = psgn(v.x) * P.Lx / 2.0 - pos.x;
        dz = psgn(v.z) * P.Lz / 2.0 - pos.z;
        dtx = dx / v.x;
        dtz = dz / v.z;
        // The reference's three cases (falling: floor; rising below the
        // ceiling: apex then floor; rising to the ceiling) each take one
        // square root.  Here the cases pick the root's argument and share
        // one sqrt, and the falling and ceiling cases share their two
        // divisions: the ceiling case's (-v.y + sqr) and (-v.y - sqr) are
        // the falling case's (sqr - v.y) and -(sqr + v.y) up to the sign of
        // a zero, which fabs drops, and its std_min operands are passed in
        // its own order.  A wave whose lanes fall and rise then runs one
        // root and two divisions instead of one set per case.
        const double y2 = v.y * v.y;
        double arg, t1 = 0.0;
        bool apex = false, ceiling = false;
        if (psgn(v.y) <= 0.0) {
            dy = pos.y + P.Ly * 0.5;
            arg = -2.0 * G_CONST * dy + y2;
        } else {
            const double maxHeight = -0.5 * y2 / G_CONST + pos.y;
            if (maxHeight < 0.5 * P.Ly) {
                apex = true;
                t1 = -v.y / G_CONST;
                const double top = pos.y + v.y * t1 + 0.5 * G_CONST * t1 * t1;
                arg = (-0.5 * P.Ly - top) * 2.0 / G_CONST;
            } else {
                ceiling = true;
                arg = y2 - 2.0 * G_CONST * (pos.y - (0.5 * P.Ly));
            }
        }
        const double sqr = sqrt(arg);
        if (apex) {
            dty = t1 + sqr;
        } else {
            const double tm = -(sqr + v.y) / G_CONST;   // ceiling case: (-v.y - sqr) / G
            const double tp = (sqr - v.y) / G_CONST;    // ceiling case: (-v.y + sqr) / G
            dty = ceiling ? std_min(fabs(tp), fabs(tm)) : std_min(fabs(tm), fabs(tp));
        }
        if (dtx < 1e-16 || __builtin_isnan(dtx))
            dtx = 1e6;
        else if (dty < 1e-16 || __builtin_isnan(dty))
            dty = 1e6;
        else if (dtz < 1e-16 || __builtin_isnan(dtz))
            dtz = 1e6;
    } else {
        dx = psgn(v.x) * P.Lx / 2.0 - pos.x;
        dy = psgn(v.y) * P.Ly / 2.0 - pos.y;
        dz = psgn(v.z) * P.Lz / 2.0 - pos.z;
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
    // Uninitialised in the reference when no branch matches (NaN times only).
    double tb = __builtin_nan("");
    char wh = K.wall_hit;
    if (dtx < dty && dtx < dtz) {
        tb = dtx; wh = 'x';
    } else if (dty < dtx && dty < dtz) {
        tb = dty; wh = 'y';
    } else if (dtz < dtx && dtz < dty) {
        tb = dtz; wh = 'z';
    } else if (dtx == dty && dtx == dtz) {
        tb = dtx; wh = 'a';
    } else if (dtx == dty) {
        tb = dtx; wh = 'b';
    } else if (dtx == dtz) {
        tb = dtx; wh = 'c';
    } else if (dty == dtz) {
        tb = dty; wh = 'd';
    }
    K.wall_hit = wh;
    const double tgas = K.next_gas - t;
    bool gas_next;
    double tcol;
    if (tgas < tb) {
        gas_next = true;
        tcol = tgas;
    } else if (tgas == tb) {
        K.next_gas += expo_draw(rng, F.tc);
        tcol = tgas;
        gas_next = false;
    } else {
        gas_next = false;
        tcol = tb;
    }
    bool reach_end;
    double tnone;
    if (tf - t <= P.maxPosStep) {
        reach_end = true;
        tnone = tf - t;
    } else {
        reach_end = false;
        tnone = P.maxPosStep;
    }
    if (tcol <= tnone) {
        K.dt = tcol;
        if (gas_next) {
            K.coll_type = 'G';
            K.n_coll += 1;
            gas_draw = true;   // K.next_gas += expo_draw(rng, F.tc), in gas_job
        } else {
            K.n_bounce += 1;
            K.coll_type = 'W';
        }
        if (tcol == tnone && reach_end) K.finished = true;
    } else {
        K.coll_type = 'N';
        K.dt = tnone;
        if (reach_end) K.finished = true;
    }
}

ST_HD inline void next_collision(const SimParams& P, const Phys& F, double tf, Kin& K, Rng& rng) {
    bool gas_draw = false;
    next_collision_core(P, F, tf, K, rng, gas_draw);
    if (gas_draw) K.next_gas += expo_draw(rng, F.tc);
}

ST_HD __forceinline__ void flip_v(d3& v, char wh) {
    // src/particle.cpp:329-371
    if (wh == 'x' || wh == 'a' || wh == 'b' || wh == 'c') v.x *= -1.0;
    if (wh == 'y' || wh == 'a' || wh == 'b' || wh == 'd') v.y *= -1.0;
    if (wh == 'z' || wh == 'a' || wh == 'c' || wh == 'd') v.z *= -1.0;
}

// The re-emissions of update_position_and_velocity, the expensive parts of
// an event, as functions of their own: move_core() names the one an event
// needs and the caller runs it (move_and_collide() below, the kinematics
// kernel's kin_step).  Running them elsewhere than in line changes nothing:
// same operations, same operands, same stream order.
enum { JOB_NONE = 0, JOB_GAS = 1, JOB_DIFFUSE = 2 };

// src/particle.cpp:306-328: Lambertian re-emission from wall wh at the
// post-flight speed Vel (v: the post-flight velocity, for the signs); the
// reference draws phi and theta before it looks at the wall id, and stops the
// particle on an edge or corner.
ST_HD inline void diffuse_job(d3& v, bool& stop, char wh, double Vel, Rng& rng) {
    const double phi = acos(sqrt(uni(rng)));
    const double theta = unif02pi(rng);
    // sincos shares one argument reduction between the two values
    // (each equals the separate sin / cos of the device libm)
    double sp, cp, sth, cth;
    sincos(phi, &sp, &cp);
    sincos(theta, &sth, &cth);
    if (wh == 'x') {
        v.x = -1.0 * psgn(v.x) * Vel * cp;
        v.y = -Vel * sp * cth;
        v.z = Vel * sp * sth;
    } else if (wh == 'y') {
        v.x = Vel * sp * cth;
        v.y = -1.0 * psgn(v.y) * Vel * cp;
        v.z = Vel * sp * sth;
    } else if (wh == 'z') {
        v.x = Vel * sp * cth;
        v.y = Vel * sp * sth;
        v.z = -1.0 * psgn(v.z) * Vel * cp;
    } else {
        stop = true;
    }
}

// A gas collision: the next gas time when it was deferred from
// next_collision_core (draw_gas_time; src/particle.cpp:228), then the
// re-emission (:373-387): 'M' gives three Maxwell-Boltzmann components, 'C'
// an isotropic direction at the current speed Vel; both draw three normals
// in the same order.
ST_HD inline void gas_job(d3& v, double& next_gas, bool draw_gas_time, double Vel, double tc, double sqrtKT_m,
                          int32_t dist, Rng& rng) {
    if (draw_gas_time) next_gas += expo_draw(rng, tc);
    if (dist == 'M' || dist == 'C') {
        double n0, n1, n2;
        normal3(rng, n0, n1, n2);
        if (dist == 'M') {
            v.x = n0 * sqrtKT_m;
            v.y = n1 * sqrtKT_m;
            v.z = n2 * sqrtKT_m;
        } else {
            const double vn = len3(d3{n0, n1, n2});
            v.x = (Vel * n0) / vn;
            v.y = (Vel * n1) / vn;
            v.z = (Vel * n2) / vn;
        }
    }
}

// src/particle.cpp:263-388 up to the re-emission: the move, the wall snap and
// the diffuse draw; returns the re-emission job the event needs and the
// post-flight speed it uses.
ST_HD inline int move_core(const SimParams& P, const Phys& F, Kin& K, Rng& rng, double& Vel) {
    K.v_old = K.v;
    K.t_old = K.t;
    K.t += K.dt;
    K.pos_old = K.pos;
    const double d = K.dt;
    const double ay = P.gravity ? G_CONST : 0.0;
    // pos = pos_old + v*dt + 0.5*a*dt*dt with a = (0, ay, 0); the x and z
    // gravity terms are (0*dt)*dt == +0 for finite dt.
    K.pos.x = (K.pos_old.x + K.v.x * d) + 0.0;
    K.pos.y = (K.pos_old.y + K.v.y * d) + ((0.5 * ay) * d) * d;
    K.pos.z = (K.pos_old.z + K.v.z * d) + 0.0;
    K.v.x = K.v.x + 0.0;
    K.v.y = K.v.y + ay * d;
    K.v.z = K.v.z + 0.0;
    Vel = len3(K.v);
    const char ct = K.coll_type, wh = K.wall_hit;
    if (ct == 'W') {
        if (wh == 'x' || wh == 'a' || wh == 'b' || wh == 'c') K.pos.x = psgn(K.pos.x) * P.Lx * 0.5;
        if (wh == 'y' || wh == 'a' || wh == 'b' || wh == 'd') K.pos.y = psgn(K.pos.y) * P.Ly * 0.5;
        if (wh == 'z' || wh == 'a' || wh == 'c' || wh == 'd') K.pos.z = psgn(K.pos.z) * P.Lz * 0.5;
        bool diffuse = false;
        if (F.diffuse_on) diffuse = uni(rng) < F.diffuse;
        if (diffuse) return JOB_DIFFUSE;
        flip_v(K.v, wh);
        return JOB_NONE;
    }
    if (ct == 'G') return JOB_GAS;
    return JOB_NONE;
}

// src/particle.cpp:263-388, after next_collision() (which drew the gas time)
ST_HD inline void move_and_collide(const SimParams& P, const Phys& F, Kin& K, Rng& rng) {
    double Vel;
    const int job = move_core(P, F, K, rng, Vel);
    if (job == JOB_DIFFUSE) diffuse_job(K.v, K.stop, K.wall_hit, Vel, rng);
    else if (job == JOB_GAS) gas_job(K.v, K.next_gas, false, Vel, F.tc, F.sqrtKT_m, F.dist, rng);
}

// src/particle.cpp:412-444
ST_HD inline void sanity(const SimParams& P, Kin& K) {
    const double tol = 1.0e-15;
    if (fabs(K.pos.x) > P.Lx / 2.0 + tol || fabs(K.pos.y) > P.Ly / 2.0 + tol ||
        fabs(K.pos.z) > P.Lz / 2.0 + tol) {
        K.stop = true; K.failure = 1; return;
    }
    if (__builtin_isinf(K.t_old) || __builtin_isinf(K.t)) { K.stop = true; K.failure = 2; return; }
    if (__builtin_isnan(K.pos_old.x) || __builtin_isnan(K.pos_old.y) || __builtin_isnan(K.pos_old.z) ||
        __builtin_isnan(K.pos.x) || __builtin_isnan(K.pos.y) || __builtin_isnan(K.pos.z)) {
        K.stop = true; K.failure = 3; return;
    }
    if (__builtin_isnan(K.v.x) || __builtin_isnan(K.v.y) || __builtin_isnan(K.v.z) ||
        __builtin_isnan(K.v_old.x) || __builtin_isnan(K.v_old.y) || __builtin_isnan(K.v_old.z)) {
        K.failure = 4; K.stop = true; return;
    }
}

// ---------------------------------------------------------------------------
// Bloch right-hand side along a ballistic segment (src/integrator.cpp:40-88)
// ---------------------------------------------------------------------------
struct Seg {
    double t0;     // segment start time (t_old)
    d3 p0, v0;     // position and velocity at t0 (pos_old, v_old)
};

// Exact: interpolate -> grad -> pulse + B0 + (1/c^2) v x E + grad -> gamma*B
// (findCrossTerm, src/integrator.cpp:63-70), in the reference's order, with
// the field hooks of spintrack_fields.h.
ST_HD __forceinline__ d3 omega_exact(const SimParams& P, const Phys& F, double t, const Seg& S) {
    const double d = t - S.t0;
    d3 p, v;
    if (P.gravity) {
        p.x = (S.p0.x + S.v0.x * d) + 0.0;
        p.y = (S.p0.y + S.v0.y * d) + (HALF_G * d) * d;
        p.z = (S.p0.z + S.v0.z * d) + 0.0;
        v.x = S.v0.x + 0.0;
        v.y = S.v0.y + G_CONST * d;
        v.z = S.v0.z + 0.0;
    } else {
        p.x = S.p0.x + S.v0.x * d;
        p.y = S.p0.y + S.v0.y * d;
        p.z = S.p0.z + S.v0.z * d;
        v = S.v0;
    }
    const double r[3] = {p.x, p.y, p.z};
    double g[3];
    st_field_grad(r, F.G, g);
    double pl[3] = {0.0, 0.0, 0.0};
    // a cos-axis pulse with a == 0 adds a signed zero, skipped bit-exactly
    // (pulse_on is set on the host when it could matter)
    if (!ST_FIELDS_PULSE_COS_AXIS || F.pulse_on) st_field_pulse(t, F.pa, F.pw, F.pulse_axis, pl);
    const double cx = v.y * F.Ez - v.z * F.Ey;
    const double cy = v.z * F.Ex - v.x * F.Ez;
    const double cz = v.x * F.Ey - v.y * F.Ex;
    const double Bx = ((pl[0] + F.B0x) + cx * IC2) + g[0];
    const double By = ((pl[1] + F.B0y) + cy * IC2) + g[1];
    const double Bz = ((pl[2] + F.B0z) + cz * IC2) + g[2];
    return d3{Bx * F.gamma, By * F.gamma, Bz * F.gamma};
}

// Bloch (:83-88): y x (gamma B)
ST_HD __forceinline__ d3 rhs_exact(const SimParams& P, const Phys& F, double t, const d3& y, const Seg& S) {
    return cross3(y, omega_exact(P, F, t, S));
}

// Fast: along a segment gamma*B is a quadratic polynomial in tau = t - t0
// (position is quadratic and velocity linear in tau, every field term is
// linear in them); the pulse, when on, is added separately.
struct Poly {
    d3 w0, w1, w2;
    // the pulse a*cos(w*t) along paxis, as gamma*a (pag) and w (pw); pon == 0:
    // no pulse term at all
    double pag, pw;
    int32_t paxis, pon;
};

// Series coefficients: row 0, cos x in x^2 ((-1)^k / (2k)!); row 1,
// sin(x)/x in x^2 ((-1)^k / (2k+1)!); row 2, atanh(s)/s in s^2
// (1 / (2k+1)); row 3, e^r in r (1 / k!).  The device reads them from
// constant memory at uniform offsets, so every Horner step is one FMA with a
// scalar-register operand: written as literals, each coefficient that is
// not an inline constant costs two extra vector moves per use (a 64-bit
// literal cannot be an FMA operand, and the accumulating form wants it in
// vector registers).
#define ST_TAYLOR_COS 1.0, -1.0 / 2, 1.0 / 24, -1.0 / 720, 1.0 / 40320, -1.0 / 3628800, 1.0 / 479001600, \
    -1.0 / 87178291200.0, 1.0 / 20922789888000.0, -1.0 / 6402373705728000.0, 1.0 / 2432902008176640000.0, 0.0, 0.0, 0.0
#define ST_TAYLOR_SINC 1.0, -1.0 / 6, 1.0 / 120, -1.0 / 5040, 1.0 / 362880, -1.0 / 39916800, 1.0 / 6227020800.0, \
    -1.0 / 1307674368000.0, 1.0 / 355687428096000.0, -1.0 / 121645100408832000.0, 1.0 / 51090942171709440000.0, \
    0.0, 0.0, 0.0
#define ST_TAYLOR_ATANH 1.0, 1.0 / 3, 1.0 / 5, 1.0 / 7, 1.0 / 9, 1.0 / 11, 1.0 / 13, 1.0 / 15, 1.0 / 17, 1.0 / 19, \
    1.0 / 21, 0.0, 0.0, 0.0
#define ST_TAYLOR_EXP 1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040, 1.0 / 40320, \
    1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0
#if defined(__HIP_DEVICE_COMPILE__)   // device-only: no host shadow to register
__constant__ double TAYLOR_DEV[4][14] = {{ST_TAYLOR_COS}, {ST_TAYLOR_SINC}, {ST_TAYLOR_ATANH}, {ST_TAYLOR_EXP}};
#endif
constexpr double TAYLOR_HOST[4][14] = {{ST_TAYLOR_COS}, {ST_TAYLOR_SINC}, {ST_TAYLOR_ATANH}, {ST_TAYLOR_EXP}};
#undef ST_TAYLOR_COS
#undef ST_TAYLOR_SINC
#undef ST_TAYLOR_ATANH
#undef ST_TAYLOR_EXP

// sum_{k<N} TAYLOR[row][k] x2^k by Horner, one FMA per term
template <int ROW, int N>
ST_HD __forceinline__ double taylor_sum(double x2) {
#pragma clang fp contract(fast)
#if defined(__HIP_DEVICE_COMPILE__)
    // The table through an opaque constant-space copy of its (uniform)
    // address: the scalar loads stay at the use (hoisted out of the lane
    // loop, the 22 coefficients would stay live in scalar registers and
    // spill), and each step is written as the three-operand FMA with the
    // coefficient in scalar registers (left to itself the compiler picks
    // the accumulating form and first copies the coefficient into vector
    // registers, two moves per step).
    typedef const __attribute__((address_space(4))) double cdouble;
    cdouble* c = (cdouble*)TAYLOR_DEV[ROW];
    asm volatile("" : "+s"(c));
    double a = c[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) {
        double r;
        asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(x2), "v"(a), "s"(c[k]));
        a = r;
    }
    return a;
#else
    const double* c = TAYLOR_HOST[ROW];
    double a = c[N - 1];
    for (int k = N - 2; k >= 0; --k) a = x2 * a + c[k];
    return a;
#endif
}

// sin and cos of a small angle, |th| <= 0.5, by their Taylor series (terms
// to th^15 and th^14; the first omitted terms are below 3e-17 relative).
ST_HD __forceinline__ void small_sincos(double th, double& s, double& c) {
#pragma clang fp contract(fast)
    const double t2 = th * th;
    c = taylor_sum<0, 8>(t2);
    s = th * taylor_sum<1, 8>(t2);
}

// Natural log and exp for the fast variant's step-size controllers, whose
// factor (tol/err)^a (prev)^b only steers the step: a few ulp suffice, at a
// third of the instructions of the library's correctly rounded forms.
// ctl_log(x), x finite and > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln x = e ln 2 + 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.1716, by the
// series to s^21 (first omitted term below 2e-17 of the sum).  ctl_exp(y),
// |y| < 700: y = n ln 2 + r (n = rint(y / ln 2), ln 2 in two parts so r is
// exact to ~1e-19), |r| <= 0.347, e^r by its Taylor series to r^13 (first
// omitted term below 5e-18), scaled by 2^n.
constexpr double LN2_HI = 6.93147180369123816490e-01;   // ln 2 to 32 bits (n * LN2_HI exact)
constexpr double LN2_LO = 1.90821492927058770002e-10;   // ln 2 - LN2_HI
constexpr double INV_LN2 = 1.44269504088896338700e+00;
ST_HD __forceinline__ double ctl_log(double x) {
#pragma clang fp contract(fast)
    if (!(x > 0.0 && x <= DBL_MAX_D)) return log(x);   // 0, inf, NaN, negative: the library's values
    int e;
    double m = frexp(x, &e);   // [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    const double r0 = __builtin_amdgcn_rcp(m + 1.0);
    const double r = __builtin_fma(r0, __builtin_fma(-(m + 1.0), r0, 1.0), r0);
    const double sv = (m - 1.0) * r;
#else
    const double sv = (m - 1.0) / (m + 1.0);
#endif
    const double p = taylor_sum<2, 11>(sv * sv);
    return (double)e * LN2_HI + ((double)e * LN2_LO + 2.0 * sv * p);
}
ST_HD __forceinline__ double ctl_exp(double y) {
#pragma clang fp contract(fast)
    // (n clamped so the conversion is defined; beyond +-2000 the result is
    // 0 or inf either way, and a NaN y gives a NaN r)
    const double n = fmin(fmax(rint(y * INV_LN2), -2000.0), 2000.0);
    const double r = (y - n * LN2_HI) - n * LN2_LO;
    return ldexp(taylor_sum<3, 14>(r), (int)n);
}

// gamma*B along a segment plus the pulse value pv (already scaled like the
// polynomial), crossed with y
template <bool PULSE = true>
ST_HD __forceinline__ d3 poly_rhs(double tau, const d3& y, const Poly& q, double pv) {
#pragma clang fp contract(fast)
    d3 w{q.w0.x + tau * (q.w1.x + tau * q.w2.x), q.w0.y + tau * (q.w1.y + tau * q.w2.y),
         q.w0.z + tau * (q.w1.z + tau * q.w2.z)};
    if (PULSE && q.pon) {
        if (q.paxis == 0) w.x += pv;
        else if (q.paxis == 1) w.y += pv;
        else w.z += pv;
    }
    return cross3(y, w);
}

ST_HD inline Poly make_poly(const SimParams& P, const Phys& F, const Seg& S) {
#pragma clang fp contract(fast)
    (void)P;
    // gamma * (B0 + G.p0 + (1/c^2) v0 x E) and gamma * (G.v0 + (1/c^2) a x E),
    // from the per-handle products of Phys (set_poly_consts)
    Poly q;
    q.w0 = d3{F.gB0[0] + F.gG[0] * S.p0.x + F.gG[1] * S.p0.y + F.gG[2] * S.p0.z + S.v0.y * F.gE[2] - S.v0.z * F.gE[1],
              F.gB0[1] + F.gG[3] * S.p0.x + F.gG[4] * S.p0.y + F.gG[5] * S.p0.z + S.v0.z * F.gE[0] - S.v0.x * F.gE[2],
              F.gB0[2] + F.gG[6] * S.p0.x + F.gG[7] * S.p0.y + F.gG[8] * S.p0.z + S.v0.x * F.gE[1] - S.v0.y * F.gE[0]};
    q.w1 = d3{F.gw1[0] + F.gG[0] * S.v0.x + F.gG[1] * S.v0.y + F.gG[2] * S.v0.z,
              F.gw1[1] + F.gG[3] * S.v0.x + F.gG[4] * S.v0.y + F.gG[5] * S.v0.z,
              F.gw1[2] + F.gG[6] * S.v0.x + F.gG[7] * S.v0.y + F.gG[8] * S.v0.z};
    q.w2 = d3{F.gw2[0], F.gw2[1], F.gw2[2]};
    q.pag = F.gpa;
    q.pw = F.pw;
    q.paxis = F.pulse_axis;
    q.pon = F.pulse_on;
    return q;
}

// The products make_poly reads: filled once per handle (or configuration)
// wherever a Phys is derived from the options.
inline void set_poly_consts(Phys& F, bool gravity) {
    const double gm = F.gamma, ay = gravity ? G_CONST : 0.0;
    for (int i = 0; i < 9; ++i) F.gG[i] = gm * F.G[i];
    F.gB0[0] = gm * F.B0x; F.gB0[1] = gm * F.B0y; F.gB0[2] = gm * F.B0z;
    F.gE[0] = gm * (F.Ex * IC2); F.gE[1] = gm * (F.Ey * IC2); F.gE[2] = gm * (F.Ez * IC2);
    // (1/c^2) a x E with a = (0, ay, 0)
    F.gw1[0] = gm * (ay * F.Ez * IC2); F.gw1[1] = 0.0; F.gw1[2] = gm * (-ay * F.Ex * IC2);
    // 0.5 * G.a
    F.gw2[0] = gm * (0.5 * ay * F.G[1]); F.gw2[1] = gm * (0.5 * ay * F.G[4]); F.gw2[2] = gm * (0.5 * ay * F.G[7]);
    F.gpa = gm * F.pa;
}

ST_HD __forceinline__ d3 rhs_fast(const SimParams& P, const Phys& F, double t, const d3& y, double t0,
                                       const Poly& q) {
#pragma clang fp contract(fast)
    return poly_rhs(t - t0, y, q, q.pon ? q.pag * cos(q.pw * t) : 0.0);
}

// gamma*B itself from the polynomial (plus the pulse when on)
ST_HD __forceinline__ d3 omega_fast(double t, double t0, const Poly& q) {
#pragma clang fp contract(fast)
    const double tau = t - t0;
    d3 w{q.w0.x + tau * (q.w1.x + tau * q.w2.x), q.w0.y + tau * (q.w1.y + tau * q.w2.y),
         q.w0.z + tau * (q.w1.z + tau * q.w2.z)};
    if (q.pon) {
        const double pv = q.pag * cos(q.pw * t);
        if (q.paxis == 0) w.x += pv;
        else if (q.paxis == 1) w.y += pv;
        else w.z += pv;
    }
    return w;
}

}  // namespace st
