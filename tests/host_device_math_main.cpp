// A CPU build of the device kinematics and RNG (spin-tracking_amd/csrc/
// st_device.h, compiled for the host) checked bit for bit against the
// oracle's restatement (oracle/st_oracle.c) on random and edge-case states:
//   * seed_rng vs or_seed (src/particle.cpp:646-663), uni / expo_draw vs
//     or_uniform / or_exponential;
//   * normal3 vs three or_normal calls (src/particle.cpp:89-107), with and
//     without a cached spare;
//   * next_collision + move_and_collide + sanity vs
//     or_calc_next_collision_time + or_update_position_and_velocity +
//     or_sanity_check (src/particle.cpp:121-444): gravity on and off, specular
//     and diffuse walls, Maxwell-Boltzmann and isotropic gas collisions, the
//     interval cap, +-0, +-inf, NaN and tiny velocity components, positions on
//     the walls, an apex exactly at the ceiling.
// The fast variant's step-size log and exp (ctl_log / ctl_exp) are checked
// against the C library to 1e-15 relative, its rotation series
// (cos_sinc_doubled) against long double cos and sin to 2e-15 absolute, and
// its rotation quaternions against the exact form (bounds below).
// Prints "cases N mismatches M" and exits non-zero on any mismatch.
//
// Diffuse re-emission takes sin and cos of two random angles: the device code
// calls sincos (the device libm's sincos returns its sin and cos), the
// oracle's gcc build merges some sin/cos pairs into glibc's sincos, which can
// differ from glibc's sin and cos in the last bit.  So the velocity after a
// diffuse wall event is compared to 1e-15 relative (the RNG stream after it
// stays bit-exact); every other output is compared bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../spin-tracking_amd/csrc/st_device.h"
#include "../spin-tracking_amd/csrc/st_integrators.h"
#include "../oracle/st_oracle.h"

static_assert(sizeof(st::Rng) == sizeof(or_rng), "rng layout");

static long g_bad = 0;
static void check(bool ok, const char* what, long i) {
    if (!ok && g_bad++ < 8) printf("mismatch: %s (case %ld)\n", what, i);
}
static bool same(double a, double b) { return memcmp(&a, &b, sizeof a) == 0; }
static bool same3(const st::d3& a, const or_coords& b) { return same(a.x, b.x) && same(a.y, b.y) && same(a.z, b.z); }
static bool close3(const st::d3& a, const or_coords& b) {
    if (same3(a, b)) return true;
    const double s = std::sqrt(b.x * b.x + b.y * b.y + b.z * b.z);
    return std::fabs(a.x - b.x) <= 1e-15 * s && std::fabs(a.y - b.y) <= 1e-15 * s && std::fabs(a.z - b.z) <= 1e-15 * s;
}
static long g_diffuse = 0, g_diffuse_inexact = 0, g_gas_jobs = 0;
static bool same_rng(const st::Rng& a, const or_rng& b) {
    return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w && same(a.spare, b.spare) && a.has == b.hasSpare;
}
static void copy_rng(st::Rng& d, const or_rng& o) {
    d.x = o.x; d.y = o.y; d.z = o.z; d.w = o.w; d.spare = o.spare; d.has = o.hasSpare;
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 400000;
    long cases = 0;
    // ---- RNG ----
    for (long i = 0; i < n; ++i) {
        or_rng o;
        or_seed(&o, 1234567u + (unsigned)(i >> 10), (unsigned)i);
        st::Rng d;
        st::seed_rng(d, 1234567u + (uint32_t)(i >> 10), (uint32_t)i);
        check(same_rng(d, o), "seed", i);
        if (i & 1) {   // leave a spare cached
            const double a = or_normal(&o, 0.0, 1.0), b = st::normal(d, 0.0, 1.0);
            check(same(a, b), "normal", i);
        }
        double r[3], g[3];
        for (int k = 0; k < 3; ++k) r[k] = or_normal(&o, 0.0, 1.0);
        st::normal3(d, g[0], g[1], g[2]);
        check(same(r[0], g[0]) && same(r[1], g[1]) && same(r[2], g[2]), "normal3", i);
        check(same(or_exponential(&o, 2e-3), st::expo_draw(d, 2e-3)), "expo", i);
        check(same(or_uniform(&o), st::uni(d)), "uniform", i);
        check(same_rng(d, o), "rng state", i);
        ++cases;
    }
    // ---- kinematics ----
    std::mt19937_64 gen(11);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    const double specials[] = {0.0, -0.0, 1e-300, -1e-300, 1e-17, -1e-17, INFINITY, -INFINITY, NAN, 1e200, -1e200,
                               1e-9, -1e-9};
    for (int grav = 0; grav < 2; ++grav)
        for (int dist = 0; dist < 2; ++dist)
            for (long i = 0; i < n; ++i) {
                or_options opt;
                memset(&opt, 0, sizeof opt);
                opt.L = or_coords{0.1, (i % 3 == 0) ? 0.02 : 0.1, 0.07};
                opt.gravity = grav;
                opt.gas_coll = true;
                opt.maxPosStep = (i % 5 == 0) ? 1e-3 : 1.0;
                opt.tc = 2e-3;
                opt.diffuse = (i % 11 == 0) ? 0.0 : 0.3;
                opt.dist = dist ? 'C' : 'M';
                opt.sqrtKT_m = 30.0;
                opt.V = 30.0;
                st::SimParams P;
                memset(&P, 0, sizeof P);
                P.Lx = opt.L.x; P.Ly = opt.L.y; P.Lz = opt.L.z;
                P.gravity = grav;
                P.gas_coll = 1;
                P.maxPosStep = opt.maxPosStep;
                st::Phys& F = P.ph;
                F.tc = opt.tc;
                F.diffuse = opt.diffuse;
                F.diffuse_on = opt.diffuse > (double)1.17549435082228750797e-38 ? 1 : 0;
                F.dist = (unsigned char)opt.dist;
                F.sqrtKT_m = opt.sqrtKT_m;
                F.V = opt.V;

                const double sc = (i % 4 == 0) ? 0.5 : 40.0;
                or_coords v{nd(gen) * sc, nd(gen) * sc, nd(gen) * sc};
                or_coords pos{(ud(gen) - 0.5) * opt.L.x, (ud(gen) - 0.5) * opt.L.y, (ud(gen) - 0.5) * opt.L.z};
                const int m = (int)(i % 23);
                if (m < 13) {
                    double* comp[] = {&v.x, &v.y, &v.z};
                    *comp[i % 3] = specials[m];
                }
                if (m == 14) pos.y = 0.5 * opt.L.y;
                if (m == 15) pos.y = -0.5 * opt.L.y;
                if (m == 16) pos.x = -0.5 * opt.L.x;
                if (m == 17 && grav)   // apex at (or a rounding step from) the ceiling
                    v.y = std::sqrt(-2.0 * st::G_CONST * (0.5 * opt.L.y - pos.y)) * (1.0 + 1e-16 * (double)(i % 3) - 1e-16);
                double t = ud(gen), next_gas = t + ud(gen) * 4e-3;
                if (m == 18) next_gas = t + 1e-12;
                const double tf = t + ((i % 7 == 0) ? 1e-4 : 0.5);

                or_rng orng;
                or_seed(&orng, 99u, (unsigned)i);
                if (i & 1) (void)or_normal(&orng, 0.0, 1.0);
                st::Rng drng;
                copy_rng(drng, orng);

                // oracle
                double o_t = t, o_told = 0.0, o_dt = 0.0, o_ng = next_gas;
                or_coords o_pos = pos, o_posold{0, 0, 0}, o_v = v, o_vold{0, 0, 0};
                char o_ct = 'N', o_wh = 'x';
                uint64_t o_nb = 0, o_nc = 0;
                bool o_fin = false, o_stop = false;
                int o_fail = 0;
                or_calc_next_collision_time(o_t, tf, o_v, o_pos, &o_ng, &o_dt, &o_ct, &o_nb, &o_nc, &o_fin, &o_wh,
                                            &orng, &opt);
                or_update_position_and_velocity(&o_told, &o_t, &o_dt, &o_posold, &o_pos, &o_vold, &o_v, &o_ct, &o_wh,
                                                &orng, &o_stop, &opt);
                or_sanity_check(&o_told, &o_t, &o_posold, &o_pos, &o_v, &o_vold, &o_stop, &o_fail, &opt);

                // device functions, host build
                st::Kin K;
                memset(&K, 0, sizeof K);
                K.t = t;
                K.next_gas = next_gas;
                K.pos = st::d3{pos.x, pos.y, pos.z};
                K.v = st::d3{v.x, v.y, v.z};
                K.coll_type = 'N';
                K.wall_hit = 'x';
                // the split order of the kinematics kernel's event (phase A,
                // then the re-emission job with the deferred gas-time draw,
                // then phase B) on a copy, which must end identical
                st::Kin K2 = K;
                st::Rng drng2 = drng;
                st::next_collision(P, F, tf, K, drng);
                st::move_and_collide(P, F, K, drng);
                st::sanity(P, K);
                {
                    bool gas_draw = false;
                    double Vel;
                    st::next_collision_core(P, F, tf, K2, drng2, gas_draw);
                    const int job = st::move_core(P, F, K2, drng2, Vel);
                    if (job == st::JOB_DIFFUSE) st::diffuse_job(K2.v, K2.stop, K2.wall_hit, Vel, drng2);
                    else if (job == st::JOB_GAS)
                        st::gas_job(K2.v, K2.next_gas, gas_draw, Vel, F.tc, F.sqrtKT_m, F.dist, drng2);
                    else if (gas_draw)
                        K2.next_gas += st::expo_draw(drng2, F.tc);
                    st::sanity(P, K2);
                    check(memcmp(&K, &K2, sizeof K) == 0 && memcmp(&drng, &drng2, sizeof drng) == 0, "split event order", i);
                    if (job == st::JOB_GAS) ++g_gas_jobs;
                }

                check(same(K.dt, o_dt) && same(K.t, o_t) && same(K.t_old, o_told) && same(K.next_gas, o_ng),
                      "times", i);
                check(same3(K.pos, o_pos) && same3(K.pos_old, o_posold), "position", i);
                // a diffuse re-emission leaves the velocity neither kept nor
                // mirrored; the specular flip only changes signs
                const bool mirrored = std::fabs(o_v.x) == std::fabs(o_vold.x) && std::fabs(o_v.z) == std::fabs(o_vold.z);
                const bool diffuse = o_ct == 'W' && F.diffuse_on && !(mirrored && (grav || std::fabs(o_v.y) == std::fabs(o_vold.y)));
                if (diffuse) {
                    ++g_diffuse;
                    if (!same3(K.v, o_v)) ++g_diffuse_inexact;
                    check(close3(K.v, o_v) && same3(K.v_old, o_vold), "velocity (diffuse)", i);
                } else {
                    check(same3(K.v, o_v) && same3(K.v_old, o_vold), "velocity", i);
                }
                check(K.coll_type == o_ct && K.wall_hit == o_wh, "collision kind", i);
                check(K.n_bounce == o_nb && K.n_coll == o_nc, "counters", i);
                check(K.finished == o_fin && K.stop == o_stop && K.failure == o_fail, "flags", i);
                check(same_rng(drng, orng), "rng after event", i);
                ++cases;
            }
    // ---- the fast controllers' log and exp (st_device.h ctl_log / ctl_exp)
    // against the C library: relative error within 1e-15 ----
    {
        std::mt19937_64 g2(5);
        std::uniform_real_distribution<double> ue(-700.0, 700.0), uy(-700.0, 700.0);
        double worst_log = 0.0, worst_exp = 0.0;
        const double xs[] = {1.0, 2.0, 0.5, 0.7071067811865476, 0.7071067811865475, 1.4142135623730951,
                             1.0000000000000002, 0.9999999999999999, 1e-300, 4.9e-324, 1.7976931348623157e308};
        const double ys[] = {0.0, -0.0, 1e-300, -1e-300, 0.5, -0.5, 0.34657359027997264, 700.0, -700.0, 709.0,
                             -708.0};
        for (long i = 0; i < n + 22; ++i) {
            const double x = i < 11 ? xs[i] : std::exp(ue(g2));
            const double y = i < 11 ? ys[i] : uy(g2);
            const double l = st::ctl_log(x), lr = std::log(x);
            const double e = st::ctl_exp(y), er = std::exp(y);
            const double dl = lr == 0.0 ? std::fabs(l) : std::fabs(l - lr) / std::fabs(lr);
            const double de = std::fabs(e - er) / er;
            worst_log = std::fmax(worst_log, dl);
            worst_exp = std::fmax(worst_exp, de);
            check(dl <= 1e-15, "ctl_log", i);
            check(de <= 1e-15, "ctl_exp", i);
            ++cases;
        }
        check(std::isnan(st::ctl_log(NAN)) && std::isinf(st::ctl_log(0.0)) && std::isinf(st::ctl_log(INFINITY)),
              "ctl_log special", 0);
        check(std::isnan(st::ctl_exp(NAN)) && st::ctl_exp(-1e6) == 0.0 && std::isinf(st::ctl_exp(1e6)),
              "ctl_exp special", 0);
        printf("ctl_log worst relative error %.3g, ctl_exp %.3g\n", worst_log, worst_exp);
    }
    // ---- the fast rotations' cos x and sin(x)/x (st_integrators.h
    // cos_sinc_doubled) in the three forms the trials use, against long
    // double cos and sin: x up to 4 (at most two extra doublings), absolute
    // error of cos x and of sin(x)/x times max(x, 1) (the rotation's
    // components) within 2e-15 ----
    {
        std::mt19937_64 g3(9);
        std::uniform_real_distribution<double> ux(-30.0, 2.0);   // log2 x
        double worst = 0.0;
        for (long i = 0; i < n; ++i) {
            const double x = std::exp2(ux(g3)), X2 = x * x;
            const long double cr = std::cos((long double)x), sr = std::sin((long double)x) / (long double)x;
            double c[3], sv[3];
            st::cos_sinc_doubled<1, false, false>(X2, c[0], sv[0]);
            st::cos_sinc_doubled<1, false, true>(X2, c[1], sv[1]);
            st::cos_sinc_doubled<0, true, false>(X2, c[2], sv[2]);
            for (int f = 0; f < 3; ++f) {
                const double ec = (double)std::fabs((long double)c[f] - cr);
                const double es = (double)std::fabs((long double)sv[f] - sr) * std::fmax(x, 1.0);
                worst = std::fmax(worst, std::fmax(ec, es));
                check(ec <= 2e-15 && es <= 2e-15, "cos_sinc_doubled", i);
            }
            ++cases;
        }
        printf("rotation series worst error %.3g\n", worst);
    }
    // ---- the fast rotation quaternions (rodriguez_quat_v<fast>, used by the
    // quaternion trials, and rot_quat_half, used by MagnusCFET) against the
    // exact form rodriguez_quat (the reference's, with the C library's sin and
    // cos): half angles p from 3e-8 to 16, absolute error of each component
    // within 1.5e-15 for p <= 2, 8e-15 for p <= 8 and 2.5e-14 above (the extra
    // doublings add a few ulp each); a zero or non-finite vector gives the
    // exact form's values (NaN components) ----
    {
        std::mt19937_64 g4(11);
        std::uniform_real_distribution<double> uu(-1.0, 1.0), lp(-25.0, 4.0);
        double worst[3] = {0.0, 0.0, 0.0};
        const double bound[3] = {1.5e-15, 8e-15, 2.5e-14};
        auto diff = [](const st::Quat& a, const st::Quat& b) {
            return std::fmax(std::fmax(std::fabs(a.w - b.w), std::fabs(a.x - b.x)),
                             std::fmax(std::fabs(a.y - b.y), std::fabs(a.z - b.z)));
        };
        for (long i = 0; i < n; ++i) {
            const st::d3 d{uu(g4), uu(g4), uu(g4)};
            const double dn = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            if (!(dn > 1e-3)) continue;
            const double p = std::exp2(lp(g4)), dt = (i & 1) ? 1.0 : 0.37, th = 2.0 * p / dt;
            const st::d3 k{d.x / dn * th, d.y / dn * th, d.z / dn * th};
            const st::Quat qe = st::rodriguez_quat(k, dt);
            const double e1 = diff(st::rodriguez_quat_v<st::VAR_FAST>(k, dt), qe);
            const double e2 = diff(st::rot_quat_half(st::d3{-0.5 * dt * k.x, -0.5 * dt * k.y, -0.5 * dt * k.z}),
                                   st::rodriguez_quat(st::d3{dt * k.x, dt * k.y, dt * k.z}, 1.0));
            const int b = p <= 2.0 ? 0 : (p <= 8.0 ? 1 : 2);
            worst[b] = std::fmax(worst[b], std::fmax(e1, e2));
            check(e1 <= bound[b] && e2 <= bound[b], "fast rotation quaternion", i);
            ++cases;
        }
        auto same_q = [](const st::Quat& a, const st::Quat& b) {
            const double x[4] = {a.w, a.x, a.y, a.z}, y[4] = {b.w, b.x, b.y, b.z};
            for (int c = 0; c < 4; ++c)
                if (std::isnan(x[c]) != std::isnan(y[c]) || (!std::isnan(x[c]) && x[c] != y[c])) return false;
            return true;
        };
        const st::d3 odd[3] = {{0.0, 0.0, 0.0}, {NAN, 0.0, 0.0}, {0.0, INFINITY, 0.0}};
        for (const st::d3& k : odd) {
            check(same_q(st::rodriguez_quat_v<st::VAR_FAST>(k, 1.0), st::rodriguez_quat(k, 1.0)), "quat_v special", 0);
            check(same_q(st::rot_quat_half(st::d3{-0.5 * k.x, -0.5 * k.y, -0.5 * k.z}), st::rodriguez_quat(k, 1.0)),
                  "rot_quat_half special", 0);
        }
        printf("rotation quaternion worst error %.3g / %.3g / %.3g\n", worst[0], worst[1], worst[2]);
    }
    printf("diffuse events %ld (last-bit differences %ld)\n", g_diffuse, g_diffuse_inexact);
    printf("gas jobs %ld\n", g_gas_jobs);
    printf("cases %ld mismatches %ld\n", cases, g_bad);
    return g_bad ? 1 : 0;
}
