// st_integrators.h -- one trial step of DOP853 / RK45 per call, so that a
// lane can interleave trial steps with event handling in one flat loop.
//
// integrateDOP (src/integrator.cpp:90-249) and integrateRK45 (:251-337)
// loop over trial steps until the segment end is reached; here the loop
// state lives in DopState / RkState and each call performs exactly one
// iteration of the reference's while(1) body.  The sequence of operations
// per iteration is the reference's.
#pragma once

#include "st_device.h"

namespace st {

// ---------------------------------------------------------------------------
// DOP853 tableau (include/coeff.h:6-176; Hairer & Wanner's DOP853)
// ---------------------------------------------------------------------------
namespace dop {
constexpr double c2 = 0.526001519587677318785587544488E-01;
constexpr double c3 = 0.789002279381515978178381316732E-01;
constexpr double c4 = 0.118350341907227396726757197510E+00;
constexpr double c5 = 0.281649658092772603273242802490E+00;
constexpr double c6 = 0.333333333333333333333333333333E+00;
constexpr double c7 = 0.25E+00;
constexpr double c8 = 0.307692307692307692307692307692E+00;
constexpr double c9 = 0.651282051282051282051282051282E+00;
constexpr double c10 = 0.6E+00;
constexpr double c11 = 0.857142857142857142857142857142E+00;
constexpr double b1 = 5.42937341165687622380535766363E-2;
constexpr double b6 = 4.45031289275240888144113950566E0;
constexpr double b7 = 1.89151789931450038304281599044E0;
constexpr double b8 = -5.8012039600105847814672114227E0;
constexpr double b9 = 3.1116436695781989440891606237E-1;
constexpr double b10 = -1.52160949662516078556178806805E-1;
constexpr double b11 = 2.01365400804030348374776537501E-1;
constexpr double b12 = 4.47106157277725905176885569043E-2;
constexpr double bhh1 = 0.244094488188976377952755905512E+00;
constexpr double bhh2 = 0.733846688281611857341361741547E+00;
constexpr double bhh3 = 0.220588235294117647058823529412E-01;
constexpr double er1 = 0.1312004499419488073250102996E-01;
constexpr double er6 = -0.1225156446376204440720569753E+01;
constexpr double er7 = -0.4957589496572501915214079952E+00;
constexpr double er8 = 0.1664377182454986536961530415E+01;
constexpr double er9 = -0.3503288487499736816886487290E+00;
constexpr double er10 = 0.3341791187130174790297318841E+00;
constexpr double er11 = 0.8192320648511571246570742613E-01;
constexpr double er12 = -0.2235530786388629525884427845E-01;
constexpr double a21 = 5.26001519587677318785587544488E-2;
constexpr double a31 = 1.97250569845378994544595329183E-2;
constexpr double a32 = 5.91751709536136983633785987549E-2;
constexpr double a41 = 2.95875854768068491816892993775E-2;
constexpr double a43 = 8.87627564304205475450678981324E-2;
constexpr double a51 = 2.41365134159266685502369798665E-1;
constexpr double a53 = -8.84549479328286085344864962717E-1;
constexpr double a54 = 9.24834003261792003115737966543E-1;
constexpr double a61 = 3.7037037037037037037037037037E-2;
constexpr double a64 = 1.70828608729473871279604482173E-1;
constexpr double a65 = 1.25467687566822425016691814123E-1;
constexpr double a71 = 3.7109375E-2;
constexpr double a74 = 1.70252211019544039314978060272E-1;
constexpr double a75 = 6.02165389804559606850219397283E-2;
constexpr double a76 = -1.7578125E-2;
constexpr double a81 = 3.70920001185047927108779319836E-2;
constexpr double a84 = 1.70383925712239993810214054705E-1;
constexpr double a85 = 1.07262030446373284651809199168E-1;
constexpr double a86 = -1.53194377486244017527936158236E-2;
constexpr double a87 = 8.27378916381402288758473766002E-3;
constexpr double a91 = 6.24110958716075717114429577812E-1;
constexpr double a94 = -3.36089262944694129406857109825E0;
constexpr double a95 = -8.68219346841726006818189891453E-1;
constexpr double a96 = 2.75920996994467083049415600797E1;
constexpr double a97 = 2.01540675504778934086186788979E1;
constexpr double a98 = -4.34898841810699588477366255144E1;
constexpr double a101 = 4.77662536438264365890433908527E-1;
constexpr double a104 = -2.48811461997166764192642586468E0;
constexpr double a105 = -5.90290826836842996371446475743E-1;
constexpr double a106 = 2.12300514481811942347288949897E1;
constexpr double a107 = 1.52792336328824235832596922938E1;
constexpr double a108 = -3.32882109689848629194453265587E1;
constexpr double a109 = -2.03312017085086261358222928593E-2;
constexpr double a111 = -9.3714243008598732571704021658E-1;
constexpr double a114 = 5.18637242884406370830023853209E0;
constexpr double a115 = 1.09143734899672957818500254654E0;
constexpr double a116 = -8.14978701074692612513997267357E0;
constexpr double a117 = -1.85200656599969598641566180701E1;
constexpr double a118 = 2.27394870993505042818970056734E1;
constexpr double a119 = 2.49360555267965238987089396762E0;
constexpr double a1110 = -3.0467644718982195003823669022E0;
constexpr double a121 = 2.27331014751653820792359768449E0;
constexpr double a124 = -1.05344954667372501984066689879E1;
constexpr double a125 = -2.00087205822486249909675718444E0;
constexpr double a126 = -1.79589318631187989172765950534E1;
constexpr double a127 = 2.79488845294199600508499808837E1;
constexpr double a128 = -2.85899827713502369474065508674E0;
constexpr double a129 = -8.87285693353062954433549289258E0;
constexpr double a1210 = 1.23605671757943030647266201528E1;
constexpr double a1211 = 6.43392746015763530355970484046E-1;
}  // namespace dop

// Loop state of one integrateDOP call (src/integrator.cpp:93-126).
struct DopState {
    double x, xf, facold, posneg;
    d3 k1;
    int32_t reject, last;
    uint32_t nstep;
};

// ---- exact variant --------------------------------------------------------
// src/integrator.cpp:114-126: call set-up, including the first RHS eval.
ST_HD __forceinline__ void dop_begin_exact(const SimParams& P, const Phys& F, DopState& D, const d3& y,
                                                const Seg& S, double t_end) {
    D.x = S.t0;
    D.xf = t_end;
    D.facold = 1.0E-4;
    D.posneg = sign_d(1.0, t_end - S.t0);
    D.last = 0;
    D.reject = 0;
    D.nstep = 0;
    D.k1 = rhs_exact(P, F, D.x, y, S);
}

#define ST_V3(dst, EX, EY, EZ) do { (dst).x = (EX); (dst).y = (EY); (dst).z = (EZ); } while (0)

// One iteration of the reference's while(1) (src/integrator.cpp:128-247).
// Returns 0 to continue, 1 when the call returns 1 (segment done), -1 when
// it returns -1 (nstep > nmax).  h is the caller's step size (tempH).
ST_HD inline int dop_trial_exact(const SimParams& P, const Phys& F, DopState& D, d3& y, double& h,
                                      const Seg& S, uint32_t& nacc, uint32_t& nrej) {
    using namespace dop;
    if (D.nstep > P.nmax) return -1;
    if ((D.x + 1.01 * h - D.xf) * D.posneg > 0.0) {
        h = D.xf - D.x;
        D.last = 1;
    }
    if (h < P.hmin && D.last != 1) h = P.hmin;
    D.nstep++;
    const double x = D.x;
    const d3 k1 = D.k1;
    d3 yy, k2, k3, k4, k5, k6, k7, k8, k9, k10;
    {
        const double ha = h * a21;
        ST_V3(yy, y.x + k1.x * ha, y.y + k1.y * ha, y.z + k1.z * ha);
    }
    k2 = rhs_exact(P, F, x + c2 * h, yy, S);
    ST_V3(yy, y.x + h * (a31 * k1.x + a32 * k2.x), y.y + h * (a31 * k1.y + a32 * k2.y),
          y.z + h * (a31 * k1.z + a32 * k2.z));
    k3 = rhs_exact(P, F, x + c3 * h, yy, S);
    ST_V3(yy, y.x + h * (a41 * k1.x + a43 * k3.x), y.y + h * (a41 * k1.y + a43 * k3.y),
          y.z + h * (a41 * k1.z + a43 * k3.z));
    k4 = rhs_exact(P, F, x + c4 * h, yy, S);
    ST_V3(yy, y.x + h * (a51 * k1.x + a53 * k3.x + a54 * k4.x),
          y.y + h * (a51 * k1.y + a53 * k3.y + a54 * k4.y),
          y.z + h * (a51 * k1.z + a53 * k3.z + a54 * k4.z));
    k5 = rhs_exact(P, F, x + c5 * h, yy, S);
    ST_V3(yy, y.x + h * (a61 * k1.x + a64 * k4.x + a65 * k5.x),
          y.y + h * (a61 * k1.y + a64 * k4.y + a65 * k5.y),
          y.z + h * (a61 * k1.z + a64 * k4.z + a65 * k5.z));
    k6 = rhs_exact(P, F, x + c6 * h, yy, S);
    ST_V3(yy, y.x + h * (a71 * k1.x + a74 * k4.x + a75 * k5.x + a76 * k6.x),
          y.y + h * (a71 * k1.y + a74 * k4.y + a75 * k5.y + a76 * k6.y),
          y.z + h * (a71 * k1.z + a74 * k4.z + a75 * k5.z + a76 * k6.z));
    k7 = rhs_exact(P, F, x + c7 * h, yy, S);
    ST_V3(yy, y.x + h * (a81 * k1.x + a84 * k4.x + a85 * k5.x + a86 * k6.x + a87 * k7.x),
          y.y + h * (a81 * k1.y + a84 * k4.y + a85 * k5.y + a86 * k6.y + a87 * k7.y),
          y.z + h * (a81 * k1.z + a84 * k4.z + a85 * k5.z + a86 * k6.z + a87 * k7.z));
    k8 = rhs_exact(P, F, x + c8 * h, yy, S);
    ST_V3(yy, y.x + h * (a91 * k1.x + a94 * k4.x + a95 * k5.x + a96 * k6.x + a97 * k7.x + a98 * k8.x),
          y.y + h * (a91 * k1.y + a94 * k4.y + a95 * k5.y + a96 * k6.y + a97 * k7.y + a98 * k8.y),
          y.z + h * (a91 * k1.z + a94 * k4.z + a95 * k5.z + a96 * k6.z + a97 * k7.z + a98 * k8.z));
    k9 = rhs_exact(P, F, x + c9 * h, yy, S);
    ST_V3(yy,
          y.x + h * (a101 * k1.x + a104 * k4.x + a105 * k5.x + a106 * k6.x + a107 * k7.x + a108 * k8.x + a109 * k9.x),
          y.y + h * (a101 * k1.y + a104 * k4.y + a105 * k5.y + a106 * k6.y + a107 * k7.y + a108 * k8.y + a109 * k9.y),
          y.z + h * (a101 * k1.z + a104 * k4.z + a105 * k5.z + a106 * k6.z + a107 * k7.z + a108 * k8.z + a109 * k9.z));
    k10 = rhs_exact(P, F, x + c10 * h, yy, S);
    ST_V3(yy,
          y.x + h * (a111 * k1.x + a114 * k4.x + a115 * k5.x + a116 * k6.x + a117 * k7.x + a118 * k8.x + a119 * k9.x + a1110 * k10.x),
          y.y + h * (a111 * k1.y + a114 * k4.y + a115 * k5.y + a116 * k6.y + a117 * k7.y + a118 * k8.y + a119 * k9.y + a1110 * k10.y),
          y.z + h * (a111 * k1.z + a114 * k4.z + a115 * k5.z + a116 * k6.z + a117 * k7.z + a118 * k8.z + a119 * k9.z + a1110 * k10.z));
    k2 = rhs_exact(P, F, x + c11 * h, yy, S);
    const double xph = x + h;
    ST_V3(yy,
          y.x + h * (a121 * k1.x + a124 * k4.x + a125 * k5.x + a126 * k6.x + a127 * k7.x + a128 * k8.x + a129 * k9.x + a1210 * k10.x + a1211 * k2.x),
          y.y + h * (a121 * k1.y + a124 * k4.y + a125 * k5.y + a126 * k6.y + a127 * k7.y + a128 * k8.y + a129 * k9.y + a1210 * k10.y + a1211 * k2.y),
          y.z + h * (a121 * k1.z + a124 * k4.z + a125 * k5.z + a126 * k6.z + a127 * k7.z + a128 * k8.z + a129 * k9.z + a1210 * k10.z + a1211 * k2.z));
    k3 = rhs_exact(P, F, xph, yy, S);
    ST_V3(k4, b1 * k1.x + b6 * k6.x + b7 * k7.x + b8 * k8.x + b9 * k9.x + b10 * k10.x + b11 * k2.x + b12 * k3.x,
          b1 * k1.y + b6 * k6.y + b7 * k7.y + b8 * k8.y + b9 * k9.y + b10 * k10.y + b11 * k2.y + b12 * k3.y,
          b1 * k1.z + b6 * k6.z + b7 * k7.z + b8 * k8.z + b9 * k9.z + b10 * k10.z + b11 * k2.z + b12 * k3.z);
    ST_V3(k5, y.x + h * k4.x, y.y + h * k4.y, y.z + h * k4.z);
    // error estimation (:191-205)
    d3 sk;
    ST_V3(sk, P.atol + P.rtol * max_d(fabs(y.x), fabs(k5.x)), P.atol + P.rtol * max_d(fabs(y.y), fabs(k5.y)),
          P.atol + P.rtol * max_d(fabs(y.z), fabs(k5.z)));
    double ex = (k4.x - bhh1 * k1.x - bhh2 * k9.x - bhh3 * k3.x) / sk.x;
    double ey = (k4.y - bhh1 * k1.y - bhh2 * k9.y - bhh3 * k3.y) / sk.y;
    double ez = (k4.z - bhh1 * k1.z - bhh2 * k9.z - bhh3 * k3.z) / sk.z;
    const double err2 = 0.0 + (ex * ex + ey * ey + ez * ez);
    ex = (er1 * k1.x + er6 * k6.x + er7 * k7.x + er8 * k8.x + er9 * k9.x + er10 * k10.x + er11 * k2.x + er12 * k3.x) / sk.x;
    ey = (er1 * k1.y + er6 * k6.y + er7 * k7.y + er8 * k8.y + er9 * k9.y + er10 * k10.y + er11 * k2.y + er12 * k3.y) / sk.y;
    ez = (er1 * k1.z + er6 * k6.z + er7 * k7.z + er8 * k8.z + er9 * k9.z + er10 * k10.z + er11 * k2.z + er12 * k3.z) / sk.z;
    double err = 0.0 + (ex * ex + ey * ey + ez * ez);
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * 3.0));
    // step-size control (:208-213)
    const double fac11 = pow(err, P.expo1);
    double fac = P.pow_beta_on ? fac11 / pow(D.facold, P.beta) : fac11 / 1.0;
    fac = max_d(P.facc2, min_d(P.facc1, fac / P.safe));
    double hnew = h / fac;
    if (err <= 1.0 || P.fixed) {
        D.facold = max_d(err, 1.0E-4);
        nacc++;
        D.k1 = rhs_exact(P, F, xph, k5, S);
        y = k5;
        D.x = xph;
        if (D.last) return 1;
        if (fabs(hnew) > P.hmax) hnew = D.posneg * P.hmax;
        if (D.reject) hnew = D.posneg * min_d(fabs(hnew), fabs(h));
        D.reject = 0;
    } else {
        nrej++;
        hnew = h / min_d(P.facc1, fac11 / P.safe);
        D.reject = 1;
        D.last = 0;
    }
    if (!P.fixed) h = hnew;
    return 0;
}

// ---- fast variant ---------------------------------------------------------
// Same trial sequence and controller.  Differences, all rounding-level:
// the RHS comes from the per-segment polynomial (rhs_fast); h is folded
// into gamma*B so each stage yields K = h*k directly; products contract into
// FMAs; the six error-norm divisions become multiplies by 1/sk; and with
// beta == 0 (the reference default) pow(err, 1/8) is three square roots.
ST_HD __forceinline__ void dop_begin_fast(const SimParams& P, const Phys& F, DopState& D, const d3& y,
                                               const Seg& S, const Poly& q, double t_end) {
    D.x = S.t0;
    D.xf = t_end;
    D.facold = 1.0E-4;
    // sign(1, t_end - t0): a segment never runs backwards (t_end >= t0)
    D.posneg = 1.0;
    D.last = 0;
    D.reject = 0;
    D.nstep = 0;
    D.k1 = rhs_fast(P, F, D.x, y, S.t0, q);
}


// Step-size arithmetic of the fast variant.  The error norm's scale factors,
// the norm's square root and err^(1/8) only steer the controller, so they use
// the hardware reciprocal / reciprocal square root refined by one Newton step
// (about 1 ulp) instead of correctly rounded division and sqrt (about 10 and
// 20 instructions each).
__device__ __forceinline__ double fast_rcp(double d) {
    const double r = __builtin_amdgcn_rcp(d);
    return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}
__device__ __forceinline__ double fast_rsq(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    // r * (1.5 - 0.5 * x * r^2)
    const double hx = 0.5 * x;
    return r * __builtin_fma(-hx * r, r, 1.5);
}
// x^(1/2) for x > 0
__device__ __forceinline__ double fast_sqrt(double x) { return x * fast_rsq(x); }

// Quadrature carried along an adjoint (backward) segment: moments
// N_n = int u(s) s^n ds, n = 0, 1, 2, of the integrated vector over the
// segment's own time s, by the step's 8th-order weights (b1, b6..b12) at the
// stage inputs; and the physical time the pulse is evaluated at,
// t = t_end - s, with the field negated (du/ds = -u x omega(t_end - s)).
struct AdjQuad {
    d3 n0, n1, n2;
    double t_end;
};


// PULSE = false compiles the pulse term out (configurations without one).
template <bool ADJ, bool PULSE = true>
ST_HD inline int dop_trial_fast_t(const SimParams& P, const Phys& F, DopState& D, d3& y, double& h,
                                  const Seg& S, const Poly& q, uint32_t& nacc, uint32_t& nrej, AdjQuad* Q) {
#pragma clang fp contract(fast)
    using namespace dop;
    if (D.nstep > P.nmax) return -1;
    if ((D.x + 1.01 * h - D.xf) * D.posneg > 0.0) {
        h = D.xf - D.x;
        D.last = 1;
    }
    if (h < P.hmin && D.last != 1) h = P.hmin;
    D.nstep++;
    const double x = D.x;
    const double tx = x - S.t0;
    // Stages yield K = h*k.  Without a pulse, the field polynomial is scaled
    // by h once per trial (hq); with one, the trial scales each k instead (3
    // multiplies per stage), which keeps the 18 registers of the scaled copy
    // free for the pulse's own state.
    Poly hq = q;
    if constexpr (!PULSE)
        hq = Poly{d3{h * q.w0.x, h * q.w0.y, h * q.w0.z}, d3{h * q.w1.x, h * q.w1.y, h * q.w1.z},
                  d3{h * q.w2.x, h * q.w2.y, h * q.w2.z}, q.pag, q.pw, q.paxis, q.pon};
    auto hk = [&](const d3& k) -> d3 {
        if constexpr (PULSE) return d3{h * k.x, h * k.y, h * k.z};
        else return k;
    };
    const d3 K1{h * D.k1.x, h * D.k1.y, h * D.k1.z};
    // The pulse a*cos(w*t) at the stage times t = x + c*h (adjoint: at
    // t_end - (x + c*h), with the field negated): one sincos per trial at the
    // step start, then cos(phi0 +- w*c*h) by angle addition with the stage
    // offset's Taylor sincos.  An offset above 0.5 rad (a step too long for
    // the pulse, which the controller rejects) is halved until it is below,
    // and its sine and cosine doubled back (a few ulp per doubling) -- a
    // smaller register footprint than a general cos in every stage.
    double pC = 1.0, pS = 0.0, hpag = 0.0;
    const bool pon = PULSE && q.pon;
    if (pon) {
        hpag = ADJ ? -q.pag : q.pag;
        sincos(q.pw * (ADJ ? (Q->t_end - x) : x), &pS, &pC);
    }
    auto pcos = [&](double c) -> double {
        double a = q.pw * (c * h), sn, cs;
        int halvings = 0;
        // (bounded: any finite angle is below 0.5 after 1100 halvings; a
        // non-finite one leaves the loop there and yields NaN)
        while (fabs(a) > 0.5 && halvings < 1100) {
            a *= 0.5;
            ++halvings;
        }
        small_sincos(a, sn, cs);
        for (; halvings > 0; --halvings) {
            const double s2 = 2.0 * sn * cs;
            cs = 1.0 - 2.0 * sn * sn;
            sn = s2;
        }
        return ADJ ? pC * cs + pS * sn : pC * cs - pS * sn;
    };
#define KR(C, YY) hk(poly_rhs<PULSE>(tx + (C) * h, (YY), hq, pon ? hpag * pcos(C) : 0.0))
    // adjoint quadrature of this trial: sum_i b_i Y_i s_i^n
    d3 q0{0.0, 0.0, 0.0}, q1{0.0, 0.0, 0.0}, q2{0.0, 0.0, 0.0};
    auto quad = [&](double b, double si, const d3& Y) {
        if constexpr (ADJ) {
            const double w1 = b * si, w2 = w1 * si;
            q0 = d3{q0.x + b * Y.x, q0.y + b * Y.y, q0.z + b * Y.z};
            q1 = d3{q1.x + w1 * Y.x, q1.y + w1 * Y.y, q1.z + w1 * Y.z};
            q2 = d3{q2.x + w2 * Y.x, q2.y + w2 * Y.y, q2.z + w2 * Y.z};
        }
    };
    quad(b1, tx, y);
    d3 yy, K2, K3, K4, K5, K6, K7, K8, K9, K10, K11, K12;
    ST_V3(yy, y.x + a21 * K1.x, y.y + a21 * K1.y, y.z + a21 * K1.z);
    K2 = KR(c2, yy);
    ST_V3(yy, y.x + (a31 * K1.x + a32 * K2.x), y.y + (a31 * K1.y + a32 * K2.y),
          y.z + (a31 * K1.z + a32 * K2.z));
    K3 = KR(c3, yy);
    ST_V3(yy, y.x + (a41 * K1.x + a43 * K3.x), y.y + (a41 * K1.y + a43 * K3.y),
          y.z + (a41 * K1.z + a43 * K3.z));
    K4 = KR(c4, yy);
    ST_V3(yy, y.x + (a51 * K1.x + a53 * K3.x + a54 * K4.x), y.y + (a51 * K1.y + a53 * K3.y + a54 * K4.y),
          y.z + (a51 * K1.z + a53 * K3.z + a54 * K4.z));
    K5 = KR(c5, yy);
    ST_V3(yy, y.x + (a61 * K1.x + a64 * K4.x + a65 * K5.x), y.y + (a61 * K1.y + a64 * K4.y + a65 * K5.y),
          y.z + (a61 * K1.z + a64 * K4.z + a65 * K5.z));
    quad(b6, tx + c6 * h, yy);
    K6 = KR(c6, yy);
    ST_V3(yy, y.x + (a71 * K1.x + a74 * K4.x + a75 * K5.x + a76 * K6.x),
          y.y + (a71 * K1.y + a74 * K4.y + a75 * K5.y + a76 * K6.y),
          y.z + (a71 * K1.z + a74 * K4.z + a75 * K5.z + a76 * K6.z));
    quad(b7, tx + c7 * h, yy);
    K7 = KR(c7, yy);
    ST_V3(yy, y.x + (a81 * K1.x + a84 * K4.x + a85 * K5.x + a86 * K6.x + a87 * K7.x),
          y.y + (a81 * K1.y + a84 * K4.y + a85 * K5.y + a86 * K6.y + a87 * K7.y),
          y.z + (a81 * K1.z + a84 * K4.z + a85 * K5.z + a86 * K6.z + a87 * K7.z));
    quad(b8, tx + c8 * h, yy);
    K8 = KR(c8, yy);
    ST_V3(yy, y.x + (a91 * K1.x + a94 * K4.x + a95 * K5.x + a96 * K6.x + a97 * K7.x + a98 * K8.x),
          y.y + (a91 * K1.y + a94 * K4.y + a95 * K5.y + a96 * K6.y + a97 * K7.y + a98 * K8.y),
          y.z + (a91 * K1.z + a94 * K4.z + a95 * K5.z + a96 * K6.z + a97 * K7.z + a98 * K8.z));
    quad(b9, tx + c9 * h, yy);
    K9 = KR(c9, yy);
    ST_V3(yy, y.x + (a101 * K1.x + a104 * K4.x + a105 * K5.x + a106 * K6.x + a107 * K7.x + a108 * K8.x + a109 * K9.x),
          y.y + (a101 * K1.y + a104 * K4.y + a105 * K5.y + a106 * K6.y + a107 * K7.y + a108 * K8.y + a109 * K9.y),
          y.z + (a101 * K1.z + a104 * K4.z + a105 * K5.z + a106 * K6.z + a107 * K7.z + a108 * K8.z + a109 * K9.z));
    quad(b10, tx + c10 * h, yy);
    K10 = KR(c10, yy);
    ST_V3(yy,
          y.x + (a111 * K1.x + a114 * K4.x + a115 * K5.x + a116 * K6.x + a117 * K7.x + a118 * K8.x + a119 * K9.x + a1110 * K10.x),
          y.y + (a111 * K1.y + a114 * K4.y + a115 * K5.y + a116 * K6.y + a117 * K7.y + a118 * K8.y + a119 * K9.y + a1110 * K10.y),
          y.z + (a111 * K1.z + a114 * K4.z + a115 * K5.z + a116 * K6.z + a117 * K7.z + a118 * K8.z + a119 * K9.z + a1110 * K10.z));
    quad(b11, tx + c11 * h, yy);
    K11 = KR(c11, yy);
    const double xph = x + h;
    ST_V3(yy,
          y.x + (a121 * K1.x + a124 * K4.x + a125 * K5.x + a126 * K6.x + a127 * K7.x + a128 * K8.x + a129 * K9.x + a1210 * K10.x + a1211 * K11.x),
          y.y + (a121 * K1.y + a124 * K4.y + a125 * K5.y + a126 * K6.y + a127 * K7.y + a128 * K8.y + a129 * K9.y + a1210 * K10.y + a1211 * K11.y),
          y.z + (a121 * K1.z + a124 * K4.z + a125 * K5.z + a126 * K6.z + a127 * K7.z + a128 * K8.z + a129 * K9.z + a1210 * K10.z + a1211 * K11.z));
    quad(b12, tx + h, yy);
    const double c12 = pon ? pcos(1.0) : 0.0;
    K12 = hk(poly_rhs<PULSE>(tx + h, yy, hq, hpag * c12));
#undef KR
    d3 bs;
    ST_V3(bs, b1 * K1.x + b6 * K6.x + b7 * K7.x + b8 * K8.x + b9 * K9.x + b10 * K10.x + b11 * K11.x + b12 * K12.x,
          b1 * K1.y + b6 * K6.y + b7 * K7.y + b8 * K8.y + b9 * K9.y + b10 * K10.y + b11 * K11.y + b12 * K12.y,
          b1 * K1.z + b6 * K6.z + b7 * K7.z + b8 * K8.z + b9 * K9.z + b10 * K10.z + b11 * K11.z + b12 * K12.z);
    d3 y1;
    ST_V3(y1, y.x + bs.x, y.y + bs.y, y.z + bs.z);
    // error norm: with K = h*k the reference's |h| * err / sqrt(n*deno)
    // becomes errK / sqrt(n*denoK)
    const double isx = fast_rcp(P.atol + P.rtol * __builtin_fmax(fabs(y.x), fabs(y1.x)));
    const double isy = fast_rcp(P.atol + P.rtol * __builtin_fmax(fabs(y.y), fabs(y1.y)));
    const double isz = fast_rcp(P.atol + P.rtol * __builtin_fmax(fabs(y.z), fabs(y1.z)));
    double ex = (bs.x - bhh1 * K1.x - bhh2 * K9.x - bhh3 * K12.x) * isx;
    double ey = (bs.y - bhh1 * K1.y - bhh2 * K9.y - bhh3 * K12.y) * isy;
    double ez = (bs.z - bhh1 * K1.z - bhh2 * K9.z - bhh3 * K12.z) * isz;
    const double err2 = ex * ex + ey * ey + ez * ez;
    ex = (er1 * K1.x + er6 * K6.x + er7 * K7.x + er8 * K8.x + er9 * K9.x + er10 * K10.x + er11 * K11.x + er12 * K12.x) * isx;
    ey = (er1 * K1.y + er6 * K6.y + er7 * K7.y + er8 * K8.y + er9 * K9.y + er10 * K10.y + er11 * K11.y + er12 * K12.y) * isy;
    ez = (er1 * K1.z + er6 * K6.z + er7 * K7.z + er8 * K8.z + er9 * K9.z + er10 * K10.z + er11 * K11.z + er12 * K12.z) * isz;
    const double e1 = ex * ex + ey * ey + ez * ez;
    const double deno = e1 + 0.01 * err2;
    const double err = (deno <= 0.0) ? 0.0 : e1 * fast_rsq(3.0 * deno);
    double fac11;
    if (!P.pow_beta_on && P.expo1 == 0.125)
        fac11 = (err > 0.0) ? fast_sqrt(fast_sqrt(fast_sqrt(err))) : 0.0;
    else
        fac11 = pow(err, P.expo1);
    double fac = P.pow_beta_on ? fac11 / pow(D.facold, P.beta) : fac11;
    fac = __builtin_fmax(P.facc2, __builtin_fmin(P.facc1, fac / P.safe));
    double hnew = h * fast_rcp(fac);
    if (err <= 1.0 || P.fixed) {
        D.facold = max_d(err, 1.0E-4);
        nacc++;
        if constexpr (ADJ) {
            // derivative at the new point for the next trial (FSAL), and the
            // accepted step's quadrature
            D.k1 = poly_rhs<PULSE>(tx + h, y1, q, -(q.pag * c12));
            Q->n0 = d3{Q->n0.x + h * q0.x, Q->n0.y + h * q0.y, Q->n0.z + h * q0.z};
            Q->n1 = d3{Q->n1.x + h * q1.x, Q->n1.y + h * q1.y, Q->n1.z + h * q1.z};
            Q->n2 = d3{Q->n2.x + h * q2.x, Q->n2.y + h * q2.y, Q->n2.z + h * q2.z};
        } else {
            D.k1 = poly_rhs<PULSE>(tx + h, y1, q, q.pag * c12);
        }
        y = y1;
        D.x = xph;
        if (D.last) return 1;
        if (fabs(hnew) > P.hmax) hnew = D.posneg * P.hmax;
        if (D.reject) hnew = D.posneg * min_d(fabs(hnew), fabs(h));
        D.reject = 0;
    } else {
        nrej++;
        hnew = h / min_d(P.facc1, fac11 / P.safe);
        D.reject = 1;
        D.last = 0;
    }
    if (!P.fixed) h = hnew;
    return 0;
}

template <bool PULSE = true>
ST_HD __forceinline__ int dop_trial_fast(const SimParams& P, const Phys& F, DopState& D, d3& y, double& h,
                                         const Seg& S, const Poly& q, uint32_t& nacc, uint32_t& nrej) {
    return dop_trial_fast_t<false, PULSE>(P, F, D, y, h, S, q, nacc, nrej, nullptr);
}

// ---------------------------------------------------------------------------
// RK45 (src/integrator.cpp:251-337, tableau include/coeff.h:178-218)
// ---------------------------------------------------------------------------
namespace rk {
constexpr double A2 = 2.0 / 9.0, A3 = 1.0 / 3.0, A4 = 3.0 / 4.0, A5 = 1.0, A6 = 5.0 / 6.0;
constexpr double B21 = 2.0 / 9.0;
constexpr double B31 = 1.0 / 12.0, B32 = 1.0 / 4.0;
constexpr double B41 = 69.0 / 128.0, B42 = -243.0 / 128.0, B43 = 135.0 / 64.0;
constexpr double B51 = -17.0 / 12.0, B52 = 27.0 / 4.0, B53 = -27.0 / 5.0, B54 = 16.0 / 15.0;
constexpr double B61 = 65.0 / 432.0, B62 = -5.0 / 16.0, B63 = 13.0 / 16.0, B64 = 4.0 / 27.0,
                 B65 = 5.0 / 144.0;
constexpr double CH1 = 47.0 / 450.0, CH2 = 0.0, CH3 = 12.0 / 25.0, CH4 = 32.0 / 225.0,
                 CH5 = 1.0 / 30.0, CH6 = 6.0 / 25.0;
constexpr double CT1 = 1.0 / 150.0, CT2 = 0.0, CT3 = -3.0 / 100.0, CT4 = 16.0 / 75.0,
                 CT5 = 1.0 / 20.0, CT6 = -6.0 / 25.0;
}  // namespace rk

// Loop state of one integrateRK45 call (src/integrator.cpp:255-266).  yy1
// persists across trials: the first trial's k1 is evaluated at the zero
// vector and later trials' k1 at the previous trial's sixth-stage input
// (:298), exactly as the reference does.
struct RkState {
    double t, tf, lastH, prev_ratio;
    double prev_lr;      // fast variant: ln(prev_ratio)
    d3 yy1;
    int32_t nstep;
    bool stop;
};

ST_HD __forceinline__ void rk_begin(RkState& R, const Seg& S, double t_end, double h) {
    R.t = S.t0;
    R.tf = t_end;
    R.stop = false;
    R.nstep = 0;
    R.lastH = h;
    R.yy1 = d3{0.0, 0.0, 0.0};
    R.prev_ratio = 1.0;
    R.prev_lr = 0.0;
}

// One iteration of the reference's while(1) (:267-335).  Returns 0 to
// continue, 1 when the call returns 0 (done), -1 / -2 for the reference's
// error returns.
template <int VAR>
ST_HD inline int rk_trial(const SimParams& P, const Phys& Fp, RkState& R, d3& y, double& h, const Seg& S,
                               const Poly& q, uint32_t& nacc, uint32_t& nrej) {
    using namespace rk;
    auto F = [&](double tt, const d3& yy) -> d3 {
        if (VAR == VAR_EXACT) return rhs_exact(P, Fp, tt, yy, S);
        return rhs_fast(P, Fp, tt, yy, S.t0, q);
    };
    R.nstep++;
    if ((uint32_t)R.nstep > P.nmax) return -2;
    const double endDt = R.tf - R.t;
    R.lastH = h;
    if (P.fixed) {
        h = P.h0;
    } else {
        if (h > P.hmax) {
            h = std_min(h, P.hmax);
        } else if (h < P.hmin) {
            return -1;
        }
    }
    if (h >= endDt) {
        h = endDt;
        R.stop = true;
    } else {
        R.lastH = h;
        R.stop = false;
    }
    const double t = R.t;
    d3 yy1 = R.yy1;
    const d3 k1 = F(t, yy1);
    {
        const double c = h * B21;
        ST_V3(yy1, y.x + c * k1.x, y.y + c * k1.y, y.z + c * k1.z);
    }
    const d3 k2 = F(t + A2 * h, yy1);
    {
        const double c1 = h * B31, c2_ = h * B32;
        ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x, y.y + c1 * k1.y + c2_ * k2.y, y.z + c1 * k1.z + c2_ * k2.z);
    }
    const d3 k3 = F(t + A3 * h, yy1);
    {
        const double c1 = h * B41, c2_ = h * B42, c3_ = h * B43;
        ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x + c3_ * k3.x, y.y + c1 * k1.y + c2_ * k2.y + c3_ * k3.y,
              y.z + c1 * k1.z + c2_ * k2.z + c3_ * k3.z);
    }
    const d3 k4 = F(t + A4 * h, yy1);
    {
        const double c1 = h * B51, c2_ = h * B52, c3_ = h * B53, c4_ = h * B54;
        ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x + c3_ * k3.x + c4_ * k4.x,
              y.y + c1 * k1.y + c2_ * k2.y + c3_ * k3.y + c4_ * k4.y,
              y.z + c1 * k1.z + c2_ * k2.z + c3_ * k3.z + c4_ * k4.z);
    }
    const d3 k5 = F(t + A5 * h, yy1);
    {
        const double c1 = h * B61, c2_ = h * B62, c3_ = h * B63, c4_ = h * B64, c5_ = h * B65;
        ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x + c3_ * k3.x + c4_ * k4.x + c5_ * k5.x,
              y.y + c1 * k1.y + c2_ * k2.y + c3_ * k3.y + c4_ * k4.y + c5_ * k5.y,
              y.z + c1 * k1.z + c2_ * k2.z + c3_ * k3.z + c4_ * k4.z + c5_ * k5.z);
    }
    const d3 k6 = F(t + A6 * h, yy1);
    R.yy1 = yy1;
    d3 ws, te;
    ST_V3(ws, y.x + h * (k1.x * CH1 + k2.x * CH2 + k3.x * CH3 + k4.x * CH4 + k5.x * CH5 + k6.x * CH6),
          y.y + h * (k1.y * CH1 + k2.y * CH2 + k3.y * CH3 + k4.y * CH4 + k5.y * CH5 + k6.y * CH6),
          y.z + h * (k1.z * CH1 + k2.z * CH2 + k3.z * CH3 + k4.z * CH4 + k5.z * CH5 + k6.z * CH6));
    ST_V3(te, h * (CT1 * k1.x + CT2 * k2.x + CT3 * k3.x + CT4 * k4.x + CT5 * k5.x + CT6 * k6.x),
          h * (CT1 * k1.y + CT2 * k2.y + CT3 * k3.y + CT4 * k4.y + CT5 * k5.y + CT6 * k6.y),
          h * (CT1 * k1.z + CT2 * k2.z + CT3 * k3.z + CT4 * k4.z + CT5 * k5.z + CT6 * k6.z));
    const double tol = P.rtol;
    double q_;
    bool small;   // error < 2 tol
    if constexpr (VAR == VAR_FAST) {
        // ratio^0.1 * prev^(-0.4/7) as one exp of a sum of logs, the previous
        // trial's log carried over, from the squared norm (no square root,
        // no division): one ctl_log and one ctl_exp instead of two pows
        const double e2 = std_max(te.x * te.x + te.y * te.y + te.z * te.z, 1.0E-32);
        const double lr = 0.5 * ctl_log(tol * tol * fast_rcp(e2));
        q_ = ctl_exp((0.7 / 7.0) * lr + (-0.4 / 7.0) * R.prev_lr);
        R.prev_lr = lr;
        small = e2 < 4.0 * tol * tol;
    } else {
        double error = len3(te);
        error = std_max(error, 1.0E-16);
        const double ratio = tol / error;
        q_ = pow(ratio, 0.7 / 7.0) * pow(R.prev_ratio, -0.4 / 7.0);
        small = error < 2 * tol;
        R.prev_ratio = ratio;
    }
    q_ = std_min(q_, 4.0);
    if ((q_ > 0.81 && small) || P.fixed) {
        nacc++;
        y = ws;
        R.t += h;
        if (R.stop) {
            h = R.lastH;
            return 1;
        }
    } else {
        nrej++;
        if (R.stop) R.stop = false;
    }
    if (!P.fixed) h = h * q_;
    return 0;
}

// ---------------------------------------------------------------------------
// integratorTypes 2-5: integrateMagnusCFET (src/integrator.cpp:620-723),
// integrateRK45Quaternion (:428-522), integrateRKF45 (:339-426) and
// integrateRKF45Quaternion (:524-618).  Same protocol as rk_trial: one call is
// one iteration of the reference's while(1), state in RkState (yy1 is used by
// RKF45 only).  The arithmetic follows the reference's order in both variants;
// the fast variant takes gamma*B from the segment polynomial.
// ---------------------------------------------------------------------------
namespace rkf {   // include/coeff.h:220-256 (Fehlberg)
constexpr double A2 = 1.0 / 4.0, A3 = 3.0 / 8.0, A4 = 12.0 / 13.0, A5 = 1.0, A6 = 0.5;
constexpr double B21 = 0.25;
constexpr double B31 = 3.0 / 32.0, B32 = 9.0 / 32.0;
constexpr double B41 = 1932.0 / 2197.0, B42 = -7200.0 / 2197.0, B43 = 7296.0 / 2197.0;
constexpr double B51 = 439.0 / 216.0, B52 = -8.0, B53 = 3680.0 / 513.0, B54 = -845.0 / 4104.0;
constexpr double B61 = -8.0 / 27.0, B62 = 2.0, B63 = -3544.0 / 2565.0, B64 = 1859.0 / 4104.0,
                 B65 = -11.0 / 40.0;
// fifth-order weights (the reference's C1..C6)
constexpr double W1 = 16.0 / 135.0, W2 = 0.0, W3 = 6656.0 / 12825.0, W4 = 28561.0 / 56430.0,
                 W5 = -9.0 / 50.0, W6 = 2.0 / 55.0;
// the fourth-order weights, which the reference's error uses directly (:398)
constexpr double R1 = 25.0 / 216.0, R2 = 0.0, R3 = 1408.0 / 2565.0, R4 = 2197.0 / 4104.0,
                 R5 = -0.2, R6 = 0.0;
}  // namespace rkf

// Quaternions (src/quaternion.cpp:128-164)
struct Quat {
    double w, x, y, z;
};

ST_HD __forceinline__ Quat q_mult(const Quat& a, const Quat& b) {
    return Quat{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}

// q v q* (:137-148)
ST_HD __forceinline__ d3 qv_mult(const Quat& q, const d3& v) {
    const Quat r = q_mult(q_mult(q, Quat{0.0, v.x, v.y, v.z}), Quat{q.w, -q.x, -q.y, -q.z});
    return d3{r.x, r.y, r.z};
}

// rotation by |k| dt about -k/|k| (:150-156)
ST_HD __forceinline__ Quat rodriguez_quat(const d3& k, double dt) {
    const double angle = len3(k);
    const double nx = k.x / angle, ny = k.y / angle, nz = k.z / angle;
    const double hh = angle * dt;
    // sincos: one argument reduction for both values (each equals the
    // device libm's separate sin / cos)
    double sn, cs;
    sincos(hh / 2.0, &sn, &cs);
    const double s = -sn;
    return Quat{cs, nx * s, ny * s, nz * s};
}

// v rotated by |k| about -k/|k| (:158-163)
ST_HD __forceinline__ d3 rodriguez(const d3& k, const d3& v) {
    const double angle = len3(k);
    double s, c;
    sincos(angle, &s, &c);   // one argument reduction (values as sin, cos)
    const d3 vk = cross3(v, k);
    const double sa = s / angle;
    const double f = ((k.x * v.x + k.y * v.y + k.z * v.z) * (1.0 - c)) / (angle * angle);
    return d3{(v.x * c + vk.x * sa) + k.x * f, (v.y * c + vk.y * sa) + k.y * f, (v.z * c + vk.z * sa) + k.z * f};
}

// Fast variant of the two rotations, without square root, division or trig
// call: cos x and sin(x)/x for x^2 <= 1 from their Taylor series in x^2
// (terms to x^20 and x^18; the first omitted terms are below 1e-19 of the
// sum), taken at x / 2^d and doubled d times: for an angle 2y,
// sin(2y)/(2y) = s c, (1 - cos 2y)/(2y)^2 = s^2/2 (no cancellation) and
// cos 2y = 1 - (2y)^2 s^2/2.  A Rodrigues rotation by theta = |k| takes them
// at theta/4 and doubles twice; a quaternion takes them at its own half
// angle.  Longer steps
// (tens of radians in one exponent: a smooth field lets the controller grow
// the step to the segment's end) halve further and take the extra doublings
// in a short loop of their own, a few ulp per doubling -- in place of the
// library's argument reduction, which one such lane would otherwise impose
// on its whole wave.  Zero vectors (the reference's 0/0), squares above 2^60
// and non-finite vectors take the exact form.  Absolute errors against the
// exact form (the C library's sin and cos): cos x and sin(x)/x within 2e-15
// for x <= 4; quaternion components within 1.5e-15 for half angles up to 2,
// 8e-15 up to 8, 2.5e-14 up to 16 (each extra doubling adds a few ulp), well
// inside the integrator tolerance the fast variant is held to
// (tests/host_device_math_main.cpp checks these bounds).
// (Q: for x^2 <= 1/4 only, 8 and 8 terms, first omitted terms below 1e-18)
template <bool Q = false>
ST_HD __forceinline__ void half_series(double x2, double& c, double& s) {
    if constexpr (Q) {
        c = taylor_sum<0, 8>(x2);
        s = taylor_sum<1, 8>(x2);
    } else {
        c = taylor_sum<0, 11>(x2);
        s = taylor_sum<1, 10>(x2);
    }
}
constexpr double FAST_ROT_MAX_X2 = 0x1p60;   // at most 31 doublings
// cos x and sin(x)/x for x^2 = X2 in (0, 2^60]: the series after d >= D0
// halvings, d the fewest that bring the square to 1 (Q: to 1/4) or below
// (VOTE: the wave decides once, with a scalar branch, whether any of its
// lanes needs more than D0 halvings, so the common case pays one multiply
// for the scaling; the branch waits for the compare, which costs more than
// it saves in the MagnusCFET chain and less in the quaternion trials.
// Q trades five series terms for one more halving and doubling in the lanes
// whose square exceeds 1/4 after D0 halvings: the MagnusCFET exponents and
// the quaternion stages' half angles are mostly below it.)
template <int D0, bool VOTE, bool Q = false>
ST_HD __forceinline__ void cos_sinc_doubled(double X2, double& c, double& s) {
#pragma clang fp contract(fast)
    constexpr double SCALE = D0 == 0 ? 1.0 : D0 == 1 ? 0.25 : 0.0625;   // 4^-D0
    constexpr int QE = Q ? 2 : 0;                                         // the range is 4^-QE/2
    static_assert(D0 >= 0 && D0 <= 2, "D0");
    const bool extra = X2 > (Q ? 0.25 : 1.0) / SCALE;
#if defined(__HIP_DEVICE_COMPILE__)
    const bool wave_extra = VOTE ? __any(extra) : true;
#else
    const bool wave_extra = VOTE ? extra : true;
#endif
    int d = D0;
    double x2 = X2 * SCALE;
    if (wave_extra) {
        int E;
        (void)frexp(X2, &E);   // X2 < 2^E, so X2 / 4^d < 2^-QE once 2d >= E + QE
        d = E + QE > 2 * D0 ? (E + QE + 1) >> 1 : D0;
        x2 = ldexp(X2, -2 * d);
    }
    half_series<Q>(x2, c, s);
    auto twice = [&]() {
        const double s2 = s * s;
        s = s * c;
        c = 1.0 - (2.0 * x2) * s2;
        x2 *= 4.0;
    };
    if (wave_extra)
        for (int i = d; i > D0; --i) twice();
#pragma unroll
    for (int i = 0; i < D0; ++i) twice();
}
template <int VAR>
ST_HD __forceinline__ d3 rodriguez_v(const d3& k, const d3& v) {
    if constexpr (VAR == VAR_FAST) {
#pragma clang fp contract(fast)
        const double t2 = k.x * k.x + k.y * k.y + k.z * k.z;
        if (t2 > 0.0 && t2 <= FAST_ROT_MAX_X2) {
            // cos(theta/2) and sin(theta/2) / (theta/2) from theta/4 (or
            // below), then the last doubling to theta
            double ch, sh;
            cos_sinc_doubled<1, false>(0.25 * t2, ch, sh);
            const double sa = sh * ch;           // sin(theta) / theta
            const double f1 = 0.5 * (sh * sh);   // (1 - cos theta) / theta^2
            const double c = 1.0 - t2 * f1;      // cos theta
            const d3 vk = cross3(v, k);
            const double f = (k.x * v.x + k.y * v.y + k.z * v.z) * f1;
            return d3{v.x * c + vk.x * sa + k.x * f, v.y * c + vk.y * sa + k.y * f, v.z * c + vk.z * sa + k.z * f};
        }
    }
    return rodriguez(k, v);
}
// Fast quaternion forms for chains of rotations.  rot_quat_half(g), g = -k/2,
// is the unit quaternion of rodriguez(k, .) -- a rotation by theta = |k|
// about -k/|k|: (cos(theta/2), g sin(theta/2)/(theta/2)) -- from the same
// series and doublings as rodriguez_v (one doubling fewer: the half angle is
// what a quaternion holds).  The caller forms g instead of k (the factor -1/2
// is exact, so g.g is exactly (theta/2)^2 and g*s exactly -(k/2)*s); the
// vectors the exact form takes go through rodriguez_quat(k, 1), which keeps
// the reference's NaN for a zero or non-finite k.  A chain of rotations is
// then composed with 16-operation quaternion products and applied to the
// vector once (q v q* as v + w t + u x t, t = 2 u x v), in place of a
// 30-operation vector rotation per factor; the same rotations, rounded
// differently.
ST_HD __forceinline__ Quat rot_quat_half(const d3& g) {
#pragma clang fp contract(fast)
    const double t2 = g.x * g.x + g.y * g.y + g.z * g.z;   // (theta/2)^2
    if (t2 > 0.0 && t2 <= 0.25 * FAST_ROT_MAX_X2) {
        double ch, sh;
        cos_sinc_doubled<1, false, true>(t2, ch, sh);
        return Quat{ch, g.x * sh, g.y * sh, g.z * sh};
    }
    return rodriguez_quat(d3{-2.0 * g.x, -2.0 * g.y, -2.0 * g.z}, 1.0);
}
ST_HD __forceinline__ Quat q_mult_fast(const Quat& a, const Quat& b) {
#pragma clang fp contract(fast)
    return Quat{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
ST_HD __forceinline__ d3 qrot_fast(const Quat& q, const d3& v) {
#pragma clang fp contract(fast)
    const double tx = 2.0 * (q.y * v.z - q.z * v.y), ty = 2.0 * (q.z * v.x - q.x * v.z),
                 tz = 2.0 * (q.x * v.y - q.y * v.x);
    return d3{v.x + q.w * tx + (q.y * tz - q.z * ty), v.y + q.w * ty + (q.z * tx - q.x * tz),
              v.z + q.w * tz + (q.x * ty - q.y * tx)};
}

// the quaternion of a rotation by |k| dt: half angle p = |k| dt / 2,
// (cos p, -k (dt/2) sin(p)/p)
template <int VAR>
ST_HD __forceinline__ Quat rodriguez_quat_v(const d3& k, double dt) {
    if constexpr (VAR == VAR_FAST) {
#pragma clang fp contract(fast)
        const double hd = 0.5 * dt;
        const double p2 = (k.x * k.x + k.y * k.y + k.z * k.z) * (hd * hd);
        if (p2 > 0.0 && p2 <= FAST_ROT_MAX_X2) {
            double c, sp;
            cos_sinc_doubled<0, true, true>(p2, c, sp);
            const double s = -(hd * sp);
            return Quat{c, k.x * s, k.y * s, k.z * s};
        }
    }
    return rodriguez_quat(k, dt);
}

// Commutator-free exponential integrators of orders 8 and 6 (Alvermann &
// Fehske, J. Comput. Phys. 230 (2011) 5930), gamma*B sampled at the 5
// Gauss-Legendre nodes; include/coeff.h:258-321 (quadrature weights folded in).
namespace cfet {
constexpr double X1 = 0.046910077030668074, X2 = 0.2307653449471585, X3 = 0.5, X4 = 0.7692346550528415,
                 X5 = 0.9530899229693319;
}  // namespace cfet

// (c1 B1 + c2 B2 + c3 B3 + c4 B4 + c5 B5) * h, summed left to right
ST_HD __forceinline__ d3 cfet_exp(double c1, double c2, double c3, double c4, double c5, const d3* B, double h) {
    return d3{((((c1 * B[0].x + c2 * B[1].x) + c3 * B[2].x) + c4 * B[3].x) + c5 * B[4].x) * h,
              ((((c1 * B[0].y + c2 * B[1].y) + c3 * B[2].y) + c4 * B[3].y) + c5 * B[4].y) * h,
              ((((c1 * B[0].z + c2 * B[1].z) + c3 * B[2].z) + c4 * B[3].z) + c5 * B[4].z) * h};
}

// the coefficient rows G1..G6 (order 8) and H1..H3 (order 6)
#define ST_G1 -0.005984812750244602, 0.020217897716672675, -0.03646660716686794, 0.04245408899634467, 0.14949496424802844
#define ST_G1R 0.14949496424802844, 0.04245408899634467, -0.03646660716686794, 0.020217897716672675, -0.005984812750244602
#define ST_G2 0.02512881778391826, -0.08955337580937157, 0.19032828600393248, 0.28920194841768715, -0.03568486888016069
#define ST_G2R -0.03568486888016069, 0.28920194841768715, 0.19032828600393248, -0.08955337580937157, 0.02512881778391826
#define ST_G3 -0.04972613261435144, 0.6572496090804754, 0.11725149008224456, -0.43611911073577403, 0.18080345083145516
#define ST_G3R 0.18080345083145516, -0.43611911073577403, 0.11725149008224456, 0.6572496090804754, -0.04972613261435144
#define ST_G4 0.0464837870954061, -0.6265183532355708, -0.11144873822286155, 0.4154988625423615, -0.17224148557040517
#define ST_G4R -0.17224148557040517, 0.4154988625423615, -0.11144873822286155, -0.6265183532355708, 0.0464837870954061
#define ST_G5 0.015567390500242426, -0.0677120111619075, -0.15241336773272682, -0.08803798509100481, -0.0013284996209207986
#define ST_G5R -0.0013284996209207986, -0.08803798509100481, -0.15241336773272682, -0.0677120111619075, 0.015567390500242426
#define ST_G6 -0.034049168494873015, 0.12263276452977083, 0.2699423185170027, 0.12263276452977094, -0.03404916849487304
#define ST_H1 -0.007431519425246311, 0.022452824849980513, -0.03620393417193995, 0.0409779901800809, 0.1516046385671249
#define ST_H1R 0.1516046385671249, 0.0409779901800809, -0.03620393417193995, 0.022452824849980513, -0.007431519425246311
#define ST_H2 0.03095749686312404, -0.10116885619385878, 0.1997455955716367, 0.3010242365525433, -0.055594729593982636
#define ST_H2R -0.055594729593982636, 0.3010242365525433, 0.1997455955716367, -0.10116885619385878, 0.03095749686312404
#define ST_H3 -0.001072443882925378, -0.023971860139062367, -0.04263887835494916, -0.023971860139062388, \
              -0.0010724438829253797

// The 16 CFET exponents in the order they act (coefficients of B1..B5): the
// order-8 chain k11 .. k1 (:688-700), then the order-6 chain k5 .. k1
// (:702-708).  Read with a uniform index by a rolled loop: one copy of the
// rotation code instead of sixteen keeps the kernel's registers in check.
// (Constant memory in the device pass; a plain table in the host pass, where
// a constant-memory variable would need the device image registered.)
#define ST_CFET_ROWS                                                                                   \
    {{ST_G1R}, {ST_G2R}, {ST_G3R}, {ST_G4R}, {ST_G5R}, {ST_G6}, {ST_G5}, {ST_G4}, {ST_G3}, {ST_G2}, {ST_G1}, \
     {ST_H1R}, {ST_H2R}, {ST_H3},  {ST_H2},  {ST_H1}}
#if defined(__HIP_DEVICE_COMPILE__)
__constant__ double CFET_ROWS[16][5] = ST_CFET_ROWS;
#else
constexpr double CFET_ROWS[16][5] = ST_CFET_ROWS;
#endif

// The same rows as moments for the fast variant.  Along a segment gamma*B is
// the quadratic w(tau) = w0 + w1 tau + w2 tau^2, so an exponent
// h sum_i c_i w(tau0 + X_i h) equals exactly
//   M0 [h w(tau0)] + M1 [h^2 w'(tau0)] + M2 [h^3 w2],  Mk = sum_i c_i X_i^k:
// nine operations per exponent from three vectors formed once per trial,
// in place of five field samples and fifteen per exponent.  Entries 3..7
// keep the row itself for the pulse, which is still sampled at the nodes.
struct CfetMoments {
    double m[16][8];
};
// (every entry times `scale`: -1/2 gives rot_quat_half's argument directly)
constexpr CfetMoments cfet_moments(double scale) {
    constexpr double rows[16][5] = ST_CFET_ROWS;
    constexpr double X[5] = {cfet::X1, cfet::X2, cfet::X3, cfet::X4, cfet::X5};
    CfetMoments M{};
    for (int r = 0; r < 16; ++r) {
        for (int i = 0; i < 5; ++i) {
            M.m[r][0] += rows[r][i];
            M.m[r][1] += rows[r][i] * X[i];
            M.m[r][2] += rows[r][i] * X[i] * X[i];
            M.m[r][3 + i] = rows[r][i];
        }
        for (int j = 0; j < 8; ++j) M.m[r][j] *= scale;
    }
    return M;
}
#if defined(__HIP_DEVICE_COMPILE__)
__constant__ CfetMoments CFET_MOM = cfet_moments(-0.5);
#endif
constexpr CfetMoments CFET_MOM_HOST = cfet_moments(-0.5);

// Quaternion integrators, per stage i = 0..5: node A_i, weight W_i of the
// step, weight E_i of the comparison solution (E_5 unused); [0] RK45
// (include/coeff.h:178-218: A, CH, C), [1] RKF45 (:220-256: A, C, C).
#define ST_QUAT_TAB                                                                                            \
    {{{0.0, 47.0 / 450.0, 1.0 / 9.0}, {2.0 / 9.0, 0.0, 0.0}, {1.0 / 3.0, 12.0 / 25.0, 9.0 / 20.0},               \
      {3.0 / 4.0, 32.0 / 225.0, 16.0 / 45.0}, {1.0, 1.0 / 30.0, 1.0 / 12.0}, {5.0 / 6.0, 6.0 / 25.0, 0.0}},        \
     {{0.0, 16.0 / 135.0, 16.0 / 135.0}, {1.0 / 4.0, 0.0, 0.0}, {3.0 / 8.0, 6656.0 / 12825.0, 6656.0 / 12825.0},   \
      {12.0 / 13.0, 28561.0 / 56430.0, 28561.0 / 56430.0}, {1.0, -9.0 / 50.0, -9.0 / 50.0},                      \
      {0.5, 2.0 / 55.0, 0.0}}}
#if defined(__HIP_DEVICE_COMPILE__)
__constant__ double QUAT_TAB[2][6][3] = ST_QUAT_TAB;
#else
constexpr double QUAT_TAB[2][6][3] = ST_QUAT_TAB;
#endif
// the same table for compile-time decisions of the fast variant
struct QuatTab {
    double q[2][6][3];
};
constexpr QuatTab QUAT_TAB_C = {ST_QUAT_TAB};
#undef ST_QUAT_TAB

// Step-size bounds of the four (e.g. :355-385): the end-of-interval test
// comes first and skips the hmax/hmin checks, and a fixed step size never
// sets stop (such a call ends only at nmax, with -2, as in the reference).
ST_HD __forceinline__ int bounds_v2(const SimParams& P, RkState& R, double& h) {
    R.nstep++;
    if ((uint32_t)R.nstep > P.nmax) return -2;
    const double endDt = R.tf - R.t;
    R.lastH = h;
    if (P.fixed) {
        h = P.h0;
    } else if (h >= endDt) {
        h = endDt;
        R.stop = true;
    } else {
        R.lastH = h;
        R.stop = false;
        if (h > P.hmax) h = std_min(h, P.hmax);
        else if (h < P.hmin) return -1;
    }
    return 0;
}

// The error argument of control_v2: the norm of the error vector d (exact
// variant), or its square (fast variant, which needs no square root).
template <int VAR>
ST_HD __forceinline__ double err_arg(const d3& d) {
    if constexpr (VAR == VAR_FAST) {
#pragma clang fp contract(fast)
        return d.x * d.x + d.y * d.y + d.z * d.z;
    } else {
        return len3(d);
    }
}

// Controller and accept test (e.g. :400-424); 1 when the call ends.  The
// fast variant forms the step factor from logs of the squared norm, as in
// rk_trial.
template <int VAR>
ST_HD __forceinline__ int control_v2(const SimParams& P, RkState& R, d3& y, const d3& ynew, double& h,
                                     double err_in, uint32_t& nacc, uint32_t& nrej) {
    const double tol = P.rtol;
    double q_;
    bool small;   // error < 2 tol
    if constexpr (VAR == VAR_FAST) {
        const double e2 = std_max(err_in, 1.0E-32);
        const double lr = 0.5 * ctl_log(tol * tol * fast_rcp(e2));
        q_ = ctl_exp((0.7 / 7.0) * lr + (-0.4 / 7.0) * R.prev_lr);
        R.prev_lr = lr;
        small = e2 < 4.0 * tol * tol;
    } else {
        const double error = std_max(err_in, 1.0E-16);
        const double ratio = tol / error;
        q_ = pow(ratio, 0.7 / 7.0) * pow(R.prev_ratio, -0.4 / 7.0);
        small = error < 2 * tol;
        R.prev_ratio = ratio;
    }
    q_ = std_min(q_, 4.0);
    if ((q_ > 0.81 && small) || P.fixed) {
        nacc++;
        y = ynew;
        R.t += h;
        if (R.stop) {
            h = R.lastH;
            return 1;
        }
    } else {
        nrej++;
        if (R.stop) R.stop = false;
    }
    if (!P.fixed) h = h * q_;
    return 0;
}

// Fast form of the quaternion trials.  A stage whose weight is zero (the
// second stage of both tableaus) contributes the reference's
// rodriguez_quat(k, 0): the identity for a finite nonzero k (multiplying by
// it changes no nonzero component), NaN for a zero or non-finite k (its
// 0/0); the factor is not formed, only that test is kept.  Where
// W_i = E_i for the first five stages (RKF45) the two chains share their
// factors: the comparison product is the step's product after stage five.
template <int INTEG, typename WF>
ST_HD __forceinline__ int quat_trial_fast(const SimParams& P, RkState& R, d3& y, double& h, double t, WF&& W,
                                          uint32_t& nacc, uint32_t& nrej) {
#pragma clang fp contract(fast)
    constexpr int TI = INTEG == INT_RKF45Q ? 1 : 0;
    constexpr bool SHARED = QUAT_TAB_C.q[TI][0][1] == QUAT_TAB_C.q[TI][0][2] &&
                            QUAT_TAB_C.q[TI][1][1] == QUAT_TAB_C.q[TI][1][2] &&
                            QUAT_TAB_C.q[TI][2][1] == QUAT_TAB_C.q[TI][2][2] &&
                            QUAT_TAB_C.q[TI][3][1] == QUAT_TAB_C.q[TI][3][2] &&
                            QUAT_TAB_C.q[TI][4][1] == QUAT_TAB_C.q[TI][4][2];
    static_assert(QUAT_TAB_C.q[TI][0][1] != 0.0 && QUAT_TAB_C.q[TI][0][2] != 0.0, "first factors start the chains");
    // the second stage carries weight 0 in both chains, every later stage
    // a nonzero one (the comparison chain ends after stage five)
    static_assert(QUAT_TAB_C.q[TI][1][1] == 0.0 && QUAT_TAB_C.q[TI][1][2] == 0.0, "stage 1 has weight 0");
    static_assert(QUAT_TAB_C.q[TI][2][1] != 0.0 && QUAT_TAB_C.q[TI][3][1] != 0.0 && QUAT_TAB_C.q[TI][4][1] != 0.0 &&
                      QUAT_TAB_C.q[TI][5][1] != 0.0 && QUAT_TAB_C.q[TI][2][2] != 0.0 &&
                      QUAT_TAB_C.q[TI][3][2] != 0.0 && QUAT_TAB_C.q[TI][4][2] != 0.0,
                  "stages 2-5 have nonzero weights");
    const double(*Q)[3] = QUAT_TAB[TI];
    d3 k = W(t);
    Quat qw = rodriguez_quat_v<VAR_FAST>(k, h * Q[0][1]);
    Quat qe = SHARED ? qw : rodriguez_quat_v<VAR_FAST>(k, h * Q[0][2]);
    bool degenerate = false;   // stage 1's zero-weight factors met a zero or non-finite k
    // (two forms of the same loop: the weight tests read from the table in
    // the two-chain trial, a branch on the stage index in the shared-chain
    // one; each is the form its kernel's register allocation takes best)
#pragma unroll 1
    for (int i = 1; i < 6; ++i) {
        k = W(t + Q[i][0] * h);
        if constexpr (SHARED) {
            if (i == 1) {   // only the test of its factors remains
                const double t2 = k.x * k.x + k.y * k.y + k.z * k.z;
                degenerate = !(t2 > 0.0 && t2 <= DBL_MAX_D);
                continue;
            }
            qw = q_mult_fast(rodriguez_quat_v<VAR_FAST>(k, h * Q[i][1]), qw);
            if (i < 5) qe = qw;
        } else {
            const double ww = Q[i][1], we = i < 5 ? Q[i][2] : 1.0;
            if (ww == 0.0 || we == 0.0) {
                const double t2 = k.x * k.x + k.y * k.y + k.z * k.z;
                degenerate = degenerate || !(t2 > 0.0 && t2 <= DBL_MAX_D);
            }
            if (ww != 0.0) qw = q_mult_fast(rodriguez_quat_v<VAR_FAST>(k, h * ww), qw);
            if (i < 5 && we != 0.0) qe = q_mult_fast(rodriguez_quat_v<VAR_FAST>(k, h * we), qe);
        }
    }
    if (degenerate) {
        const double nan = __builtin_nan("");
        qw = qe = Quat{nan, nan, nan, nan};
    }
    const d3 ws = qrot_fast(qw, y);
    const d3 te = qrot_fast(qe, y);
    const d3 d{ws.x - te.x, ws.y - te.y, ws.z - te.z};
    return control_v2<VAR_FAST>(P, R, y, ws, h, err_arg<VAR_FAST>(d), nacc, nrej);
}

template <int VAR, int INTEG>
ST_HD inline int rk2_trial(const SimParams& P, const Phys& Fp, RkState& R, d3& y, double& h, const Seg& S,
                           const Poly& q, uint32_t& nacc, uint32_t& nrej) {
    auto W = [&](double tt) -> d3 {
        if (VAR == VAR_EXACT) return omega_exact(P, Fp, tt, S);
        return omega_fast(tt, S.t0, q);
    };
    {
        const int b = bounds_v2(P, R, h);
        if (b) return b;
    }
    const double t = R.t;
    if constexpr (INTEG == INT_RKF45) {
        using namespace rkf;
        auto F = [&](double tt, const d3& yy) -> d3 { return cross3(yy, W(tt)); };
        d3 yy1 = R.yy1;
        const d3 k1 = F(t, yy1);
        {
            const double c = h * B21;
            ST_V3(yy1, y.x + c * k1.x, y.y + c * k1.y, y.z + c * k1.z);
        }
        const d3 k2 = F(t + A2 * h, yy1);
        {
            const double c1 = h * B31, c2_ = h * B32;
            ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x, y.y + c1 * k1.y + c2_ * k2.y, y.z + c1 * k1.z + c2_ * k2.z);
        }
        const d3 k3 = F(t + A3 * h, yy1);
        {
            const double c1 = h * B41, c2_ = h * B42, c3_ = h * B43;
            ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x + c3_ * k3.x, y.y + c1 * k1.y + c2_ * k2.y + c3_ * k3.y,
                  y.z + c1 * k1.z + c2_ * k2.z + c3_ * k3.z);
        }
        const d3 k4 = F(t + A4 * h, yy1);
        {
            const double c1 = h * B51, c2_ = h * B52, c3_ = h * B53, c4_ = h * B54;
            ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x + c3_ * k3.x + c4_ * k4.x,
                  y.y + c1 * k1.y + c2_ * k2.y + c3_ * k3.y + c4_ * k4.y,
                  y.z + c1 * k1.z + c2_ * k2.z + c3_ * k3.z + c4_ * k4.z);
        }
        const d3 k5 = F(t + A5 * h, yy1);
        {
            const double c1 = h * B61, c2_ = h * B62, c3_ = h * B63, c4_ = h * B64, c5_ = h * B65;
            ST_V3(yy1, y.x + c1 * k1.x + c2_ * k2.x + c3_ * k3.x + c4_ * k4.x + c5_ * k5.x,
                  y.y + c1 * k1.y + c2_ * k2.y + c3_ * k3.y + c4_ * k4.y + c5_ * k5.y,
                  y.z + c1 * k1.z + c2_ * k2.z + c3_ * k3.z + c4_ * k4.z + c5_ * k5.z);
        }
        const d3 k6 = F(t + A6 * h, yy1);
        R.yy1 = yy1;
        d3 ws, te;
        ST_V3(ws, y.x + h * (k1.x * W1 + k2.x * W2 + k3.x * W3 + k4.x * W4 + k5.x * W5 + k6.x * W6),
              y.y + h * (k1.y * W1 + k2.y * W2 + k3.y * W3 + k4.y * W4 + k5.y * W5 + k6.y * W6),
              y.z + h * (k1.z * W1 + k2.z * W2 + k3.z * W3 + k4.z * W4 + k5.z * W5 + k6.z * W6));
        ST_V3(te, h * (R1 * k1.x + R2 * k2.x + R3 * k3.x + R4 * k4.x + R5 * k5.x + R6 * k6.x),
              h * (R1 * k1.y + R2 * k2.y + R3 * k3.y + R4 * k4.y + R5 * k5.y + R6 * k6.y),
              h * (R1 * k1.z + R2 * k2.z + R3 * k3.z + R4 * k4.z + R5 * k5.z + R6 * k6.z));
        return control_v2<VAR>(P, R, y, ws, h, err_arg<VAR>(te), nacc, nrej);
    } else if constexpr (INTEG == INT_RK45Q || INTEG == INT_RKF45Q) {
        // k_i = gamma*B(t + A_i h); the step rotates y by the product of
        // R(k_i, h W_i), k1 first; the comparison solution drops the sixth
        // factor and uses weights E_i (RK45: W = CH, E = C; RKF45: W = E = C)
        if constexpr (VAR == VAR_FAST) return quat_trial_fast<INTEG>(P, R, y, h, t, W, nacc, nrej);
        const double(*Q)[3] = QUAT_TAB[INTEG == INT_RKF45Q ? 1 : 0];
        d3 k = W(t);
        Quat qw = rodriguez_quat_v<VAR>(k, h * Q[0][1]);
        Quat qe = rodriguez_quat_v<VAR>(k, h * Q[0][2]);
#pragma unroll 1
        for (int i = 1; i < 6; ++i) {
            k = W(t + Q[i][0] * h);
            qw = q_mult(rodriguez_quat_v<VAR>(k, h * Q[i][1]), qw);
            if (i < 5) qe = q_mult(rodriguez_quat_v<VAR>(k, h * Q[i][2]), qe);
        }
        const d3 ws = qv_mult(qw, y);
        const d3 te = qv_mult(qe, y);
        const d3 d{ws.x - te.x, ws.y - te.y, ws.z - te.z};
        return control_v2<VAR>(P, R, y, ws, h, err_arg<VAR>(d), nacc, nrej);
    } else {
        static_assert(INTEG == INT_MAGNUS, "integrator type");
        using namespace cfet;
        if constexpr (VAR == VAR_FAST) {
#pragma clang fp contract(fast)
            // the exponents from the segment polynomial's moments (CFET_MOM,
            // scaled by -1/2 for rot_quat_half)
            const double tau = t - S.t0, h2 = h * h, h3 = h2 * h, tt = 2.0 * tau;
            const d3 u0{h * (q.w0.x + tau * (q.w1.x + tau * q.w2.x)), h * (q.w0.y + tau * (q.w1.y + tau * q.w2.y)),
                        h * (q.w0.z + tau * (q.w1.z + tau * q.w2.z))};
            const d3 u1{h2 * (q.w1.x + tt * q.w2.x), h2 * (q.w1.y + tt * q.w2.y), h2 * (q.w1.z + tt * q.w2.z)};
            const d3 u2{h3 * q.w2.x, h3 * q.w2.y, h3 * q.w2.z};
            // h * pulse at the five nodes
            double hp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            const bool pon = q.pon;
            if (pon) {
                const double X[5] = {X1, X2, X3, X4, X5};
#pragma unroll
                for (int i = 0; i < 5; ++i) hp[i] = h * (q.pag * cos(q.pw * (t + X[i] * h)));
            }
            auto expo = [&](int r) -> d3 {
#if defined(__HIP_DEVICE_COMPILE__)
                // an opaque constant-space pointer: the row's scalar loads stay
                // here (the two rows of the first factors would otherwise be
                // hoisted out of the lane loop and spilled to vector lanes)
                typedef const __attribute__((address_space(4))) double cdouble;
                cdouble* m = (cdouble*)CFET_MOM.m[r];
                asm volatile("" : "+s"(m));
#else
                const double* m = CFET_MOM_HOST.m[r];
#endif
                d3 k{m[0] * u0.x + m[1] * u1.x + m[2] * u2.x, m[0] * u0.y + m[1] * u1.y + m[2] * u2.y,
                     m[0] * u0.z + m[1] * u1.z + m[2] * u2.z};
                if (pon) {
                    const double pe = m[3] * hp[0] + m[4] * hp[1] + m[5] * hp[2] + m[6] * hp[3] + m[7] * hp[4];
                    k.x += q.paxis == 0 ? pe : 0.0;
                    k.y += q.paxis == 1 ? pe : 0.0;
                    k.z += q.paxis == 2 ? pe : 0.0;
                }
                return k;
            };
            // both chains as quaternion products (the first factor acts
            // first; each chain starts at its first factor, which equals that
            // factor times the identity exactly)
            // (two independent rotations per loop trip: the two chains side
            // by side while the order-6 chain lasts, then two consecutive
            // order-8 factors, whose products with the chain follow in order)
            Quat q8 = rot_quat_half(expo(0)), q6 = rot_quat_half(expo(11));
#pragma unroll 1
            for (int r = 1; r < 5; ++r) {
                q8 = q_mult_fast(rot_quat_half(expo(r)), q8);
                q6 = q_mult_fast(rot_quat_half(expo(11 + r)), q6);
            }
#pragma unroll 1
            for (int r = 5; r < 11; r += 2) {
                const Quat a = rot_quat_half(expo(r)), b = rot_quat_half(expo(r + 1));
                q8 = q_mult_fast(b, q_mult_fast(a, q8));
            }
            const d3 y8 = qrot_fast(q8, y), y6 = qrot_fast(q6, y);
            const d3 d{y8.x - y6.x, y8.y - y6.y, y8.z - y6.z};
            return control_v2<VAR>(P, R, y, y8, h, err_arg<VAR>(d), nacc, nrej);
        }
        const d3 B[5] = {W(t + X1 * h), W(t + X2 * h), W(t + X3 * h), W(t + X4 * h), W(t + X5 * h)};
        // order 8: k11 (row 1 reversed) acts first ... k6 (row 6) ... k1
        // (row 1); order 6: k5 (row 1 reversed), k4 (row 2 reversed), k3,
        // k2, k1 -- each chain starting from y.  The two chains are
        // independent: their first five rotations go side by side (the same
        // operations on each chain, so the same results), which shortens the
        // dependent path from 16 rotations to 11 when few waves are resident.
        d3 y8 = y, y6 = y;
#pragma unroll 1
        for (int r = 0; r < 11; ++r) {
            const double* c = CFET_ROWS[r];
            y8 = rodriguez_v<VAR>(cfet_exp(c[0], c[1], c[2], c[3], c[4], B, h), y8);
            if (r < 5) {
                const double* c6 = CFET_ROWS[11 + r];
                y6 = rodriguez_v<VAR>(cfet_exp(c6[0], c6[1], c6[2], c6[3], c6[4], B, h), y6);
            }
        }
        const d3 d{y8.x - y6.x, y8.y - y6.y, y8.z - y6.z};
        return control_v2<VAR>(P, R, y, y8, h, err_arg<VAR>(d), nacc, nrej);
    }
}

// every non-DOP853 integrator behind one call
template <int VAR, int INTEG>
ST_HD __forceinline__ int rk_any_trial(const SimParams& P, const Phys& Fp, RkState& R, d3& y, double& h,
                                       const Seg& S, const Poly& q, uint32_t& nacc, uint32_t& nrej) {
    if constexpr (INTEG == INT_RK45)
        return rk_trial<VAR>(P, Fp, R, y, h, S, q, nacc, nrej);
    else
        return rk2_trial<VAR, INTEG>(P, Fp, R, y, h, S, q, nacc, nrej);
}

// the eleven order-8 rows in the order their exponentials act (row 1
// reversed first), for the host-side propagator integration
constexpr double CFET8_ACTING[11][5] = {{ST_G1R}, {ST_G2R}, {ST_G3R}, {ST_G4R}, {ST_G5R}, {ST_G6},
                                       {ST_G5},  {ST_G4},  {ST_G3},  {ST_G2},  {ST_G1}};

#undef ST_G1
#undef ST_G1R
#undef ST_G2
#undef ST_G2R
#undef ST_G3
#undef ST_G3R
#undef ST_G4
#undef ST_G4R
#undef ST_G5
#undef ST_G5R
#undef ST_G6
#undef ST_H1
#undef ST_H1R
#undef ST_H2
#undef ST_H2R
#undef ST_H3
#undef ST_V3

}  // namespace st
