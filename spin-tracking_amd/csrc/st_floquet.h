// st_floquet.h -- host side of the spectral path (integratorType 6): the
// Floquet decomposition of the driven single-spin propagator and the
// Floquet-Markov evolution that turns the noise spectrum the GPU
// accumulates into the final Bloch vector.  One-off host work per run, as in
// the reference (floquet_diagonalize at initParticles, floquet_integrate at
// postProcess).
//
//   floquet_diagonalize            src/integrator.cpp:824-867
//   integrateHamiltonian           src/integrator.cpp:745-789
//   floquet_master_equation_rates  src/floquet.cpp:192-279
//   Spectrum::lookup               src/floquet.cpp:51-64
//   integrateFloquetMarkov         src/integrator.cpp:791-822
//   floquet_integrate              src/integrator.cpp:869-889
//   bloch_to_density / density_to_bloch   src/floquet.cpp:281-301
//   qEigenval, qEigenvec, toSU2, pow      src/quaternion.cpp:169-201
//
// Arithmetic follows the reference's operation order (the library builds
// with -ffp-contract=off and the host libm).  The 2x2 matrix exponentials
// use the closed form exp(A) = e^m (cosh(s) I + sinh(s)/s (A - m I)),
// m = tr(A)/2, s^2 = ((a-d)/2)^2 + bc (Cayley-Hamilton), where the
// reference calls Eigen's Pade scaling-and-squaring; both are accurate to
// rounding for these 2x2 rate matrices.  Spectrum::lookup past its last
// frequency reads one element beyond the reference's array
// (src/floquet.cpp:63); here it returns the last tabulated power.
#pragma once

#include <cmath>
#include <complex>
#include <vector>

#include "st_spectrum.h"

namespace st {
namespace floquet {

using cd = std::complex<double>;

struct M2 {
    cd a[2][2];
};

inline Quat qconj(const Quat& q) { return Quat{q.w, -q.x, -q.y, -q.z}; }
inline Quat qscale(const Quat& q, double a) { return Quat{q.w * a, q.x * a, q.y * a, q.z * a}; }
inline Quat qadd(const Quat& a, const Quat& b) { return Quat{a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }

// qEigenval / qEigenvec (src/quaternion.cpp:169-177)
inline Quat q_eigenval(const Quat& q) { return Quat{q.w, 0.0, 0.0, -std::sqrt(1.0 - q.w * q.w)}; }
inline Quat q_eigenvec(const Quat& q) {
    const Quat q1{std::sqrt(1.0 - q.w * q.w) - q.z, q.y, -q.x, 0.0};
    const double n = std::sqrt(q1.w * q1.w + q1.x * q1.x + q1.y * q1.y + q1.z * q1.z);
    return qscale(q1, 1.0 / n);
}

// pow(quaternion, int) by halving (src/quaternion.cpp:190-201)
inline Quat qpow(const Quat& q, int n) {
    if (n == 0) return Quat{1.0, 0.0, 0.0, 0.0};
    if (n == 1) return q;
    const int half = n / 2;
    const Quat q2 = qpow(q, half);
    const Quat qr = qpow(q, n - half - half);
    return q_mult(q_mult(q2, q2), qr);
}

inline M2 mul(const M2& x, const M2& y) {
    M2 r;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j];
    return r;
}
inline M2 adjoint(const M2& x) {
    M2 r;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r.a[i][j] = std::conj(x.a[j][i]);
    return r;
}

// toSU2 (src/quaternion.cpp:179-188)
inline M2 to_su2(const Quat& q) {
    M2 r;
    r.a[0][0] = cd(q.w, q.z);
    r.a[0][1] = cd(q.y, q.x);
    r.a[1][0] = cd(-q.y, q.x);
    r.a[1][1] = cd(q.w, -q.z);
    return r;
}

inline M2 bloch_to_density(const d3& b, const Quat& basis) {
    const M2 B = to_su2(basis);
    M2 rho;
    rho.a[0][0] = cd(1 + b.z, 0.0);
    rho.a[0][1] = cd(b.x, -b.y);
    rho.a[1][0] = cd(b.x, b.y);
    rho.a[1][1] = cd(1 - b.z, 0.0);
    M2 r = mul(mul(B, rho), adjoint(B));
    for (auto& row : r.a)
        for (auto& v : row) v /= 2.0;
    return r;
}

inline d3 density_to_bloch(const M2& rho, const Quat& basis) {
    const M2 B = to_su2(basis);
    const M2 L = mul(mul(adjoint(B), rho), B);
    // traces of L sigma_x, L sigma_y, L sigma_z
    const cd tx = L.a[0][1] + L.a[1][0];
    const cd ty = L.a[0][1] * cd(0.0, 1.0) + L.a[1][0] * cd(0.0, -1.0);
    const cd tz = L.a[0][0] - L.a[1][1];
    return d3{tx.real(), ty.real(), tz.real()};
}

// exp(A) of a complex 2x2 matrix, closed form (see the header comment)
inline M2 expm2(const M2& A) {
    const cd m = (A.a[0][0] + A.a[1][1]) / 2.0;
    const cd dl = (A.a[0][0] - A.a[1][1]) / 2.0;
    const cd s = std::sqrt(dl * dl + A.a[0][1] * A.a[1][0]);
    cd ch, shs;   // cosh(s), sinh(s)/s
    if (std::abs(s) < 1e-4) {
        const cd s2 = s * s;
        ch = 1.0 + s2 / 2.0 * (1.0 + s2 / 12.0 * (1.0 + s2 / 30.0));
        shs = 1.0 + s2 / 6.0 * (1.0 + s2 / 20.0 * (1.0 + s2 / 42.0));
    } else {
        ch = std::cosh(s);
        shs = std::sinh(s) / s;
    }
    const cd em = std::exp(m);
    M2 r;
    r.a[0][0] = em * (ch + shs * dl);
    r.a[1][1] = em * (ch - shs * dl);
    r.a[0][1] = em * shs * A.a[0][1];
    r.a[1][0] = em * shs * A.a[1][0];
    return r;
}

// integrateHamiltonian (src/integrator.cpp:745-789): fixed steps h of the
// order-8 CFET propagator along the field at the origin at rest; y <- q1 q2
// ... q11 y.  Returns -2 after nmax steps as the reference does.
inline int integrate_hamiltonian(const SimParams& P, double t0, double tf, Quat& y, double h, uint32_t nmax) {
    const Seg zero{t0, d3{0.0, 0.0, 0.0}, d3{0.0, 0.0, 0.0}};
    double t = t0;
    uint32_t nstep = 0;
    while (t < tf) {
        nstep++;
        if (nstep > nmax) return -2;
        const double rest = tf - t;
        if (h >= rest) h = rest;
        d3 B[5];
        const double X[5] = {cfet::X1, cfet::X2, cfet::X3, cfet::X4, cfet::X5};
        for (int i = 0; i < 5; ++i) B[i] = omega_exact(P, P.ph, t + X[i] * h, zero);
        // q_k = exp of row CFET8_ACTING[11 - k]; P = ((q1 q2) q3) ... q11
        Quat prod{1.0, 0.0, 0.0, 0.0};
        for (int r = 10; r >= 0; --r) {
            const double* c = CFET8_ACTING[r];
            const Quat q = rodriguez_quat(cfet_exp(c[0], c[1], c[2], c[3], c[4], B, 1.0), h);
            prod = r == 10 ? q : q_mult(prod, q);
        }
        y = q_mult(prod, y);
        t += h;
    }
    return 0;
}

struct Diag {
    Quat f_modes_0, f_energies;
    double freq[NW];
    double dt;
    int n_prop;
    std::vector<Quat> props;
};

// floquet_diagonalize (src/integrator.cpp:824-867)
inline int diagonalize(const SimParams& P, double w, uint32_t nmax, Diag& fd) {
    const double t0 = 0.0;
    const double tf = (2 * M_PI) / w;
    const int n_prop = 100;
    fd.props.assign(n_prop, Quat{1.0, 0.0, 0.0, 0.0});
    Quat y{1.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < n_prop; ++i) {
        const double t1 = t0 + (tf - t0) * i / n_prop;
        const double t2 = t0 + (tf - t0) * (i + 1) / n_prop;
        (void)integrate_hamiltonian(P, t1, t2, y, 1e-6, nmax);   // the reference ignores the status
        fd.props[i] = y;
    }
    fd.f_energies = q_eigenval(fd.props[n_prop - 1]);
    fd.f_modes_0 = q_eigenvec(fd.props[n_prop - 1]);
    const double ea = std::fabs(std::atan2(fd.f_energies.z, fd.f_energies.w)) / (tf - t0);
    const double eb = -ea;
    const double dE = ea - eb;
    for (int k = 0; k <= NK / 2; ++k)
        for (int i = -1; i < 2; ++i) {
            const int idx = k * 3 + i;
            if (idx >= 0) fd.freq[idx] = dE * i + k * w;
        }
    fd.dt = (tf - t0) / n_prop;
    fd.n_prop = n_prop;
    return 0;
}

// Spectrum::lookup (src/floquet.cpp:51-64)
inline cd lookup(const double* freq, const cd* power, double f) {
    if (f < 0) return std::conj(lookup(freq, power, -f));
    double best = INFINITY;
    for (int i = 0; i < NW; ++i) {
        const double d = std::fabs(freq[i] - f);
        if (d < best) best = d;
        else return power[i - 1];
    }
    return power[NW - 1];
}

inline double heaviside(double x) {
    if (x == 0.0) return 0.5;
    return ((x < 0.0 ? -1.0 : 1.0) + 1) / 2;
}

// floquet_master_equation_rates (src/floquet.cpp:192-279): Delta, X, Gamma
// are overwritten, Zeta and Omicron accumulate.
inline void rates(const Quat& f_modes_0, const Quat& f_energies, const Quat& c_op, const Quat* props, int n_prop,
                  double period, const double* freq, const cd* power, double (&Delta)[2][2][NK],
                  cd (&X)[2][2][NK], cd (&Gamma)[2][2][NK], cd (&Zeta)[2][2], cd (&Om)[2][2]) {
    Quat Xq[NK];
    for (auto& q : Xq) q = Quat{0.0, 0.0, 0.0, 0.0};
    double Xsq[2][2][NK];
    const double omega = 2 * M_PI / period;
    const double ea = std::atan2(f_energies.z, f_energies.w) / period;
    const double eb = -std::atan2(f_energies.z, f_energies.w) / period;
    const double es[2] = {ea, eb};
    const int kmax = NK / 2;
    for (int i = 0; i < n_prop; ++i) {
        const double t = (i + 1) * period / n_prop;
        const double weight = 1.0 / n_prop;
        const Quat phase{std::cos(ea * t), 0.0, 0.0, -std::sin(ea * t)};
        const Quat fm = q_mult(q_mult(props[i], f_modes_0), phase);
        const Quat q = q_mult(q_mult(qconj(fm), c_op), fm);
        for (int k = 0; k < NK; ++k) {
            const double arg = (k - kmax) * omega * t;
            const Quat kp{std::cos(arg), 0.0, 0.0, std::sin(arg)};
            Xq[k] = qadd(Xq[k], qscale(q_mult(q, kp), weight));
        }
    }
    for (int k = 0; k <= kmax * 2; ++k) {
        X[0][0][k] = cd(Xq[k].z, -Xq[k].w);
        X[1][1][k] = std::conj(X[0][0][k]);
        X[0][1][k] = cd(Xq[k].x, -Xq[k].y);
        X[1][0][k] = cd(Xq[2 * kmax - k].x, Xq[2 * kmax - k].y);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) Xsq[i][j][k] = std::norm(X[i][j][k]);
    }
    for (int k = 0; k < NK; ++k)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                const double f = (es[j] - es[i]) + (k - NK / 2) * omega;
                Delta[i][j][k] = f;
                Gamma[i][j][k] = Xsq[i][j][k] * lookup(freq, power, f) * 2.0 * M_PI * heaviside(f);
            }
    for (int k = 0; k <= kmax * 2; ++k)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) Zeta[i][j] = Zeta[i][j] + Xsq[i][j][k] * lookup(freq, power, Delta[i][j][k]) / 2.0;
    for (int k = 0; k <= kmax * 2; ++k)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                const double f = (k - NK / 2) * omega;
                Om[i][j] = Om[i][j] + lookup(freq, power, f) * X[i][i][k] * std::conj(X[j][j][k]);
            }
}

// integrateFloquetMarkov (src/integrator.cpp:791-822)
inline void markov(double t0, double tf, M2& rho, const cd (&Z)[2][2], const cd (&Om)[2][2]) {
    const double dt = tf - t0;
    M2 Ad;
    Ad.a[0][0] = (-Z[1][0] + Z[0][1]) * dt;
    Ad.a[0][1] = (Z[0][1] + Z[1][0]) * dt;
    Ad.a[1][0] = (Z[1][0] + Z[0][1]) * dt;
    Ad.a[1][1] = (-Z[1][0] + Z[0][1]) * dt;
    const M2 Ed = expm2(Ad);
    const cd p0 = rho.a[0][0], p1 = rho.a[1][1];
    rho.a[0][0] = Ed.a[0][0] * p0 + Ed.a[0][1] * p1;
    rho.a[1][1] = Ed.a[1][0] * p0 + Ed.a[1][1] * p1;
    const cd d01 = -(Z[0][0] + Z[0][1] + Z[0][1] + Z[1][1]);
    const cd d10 = -(Z[1][0] + Z[1][1] + Z[0][0] + Z[1][0]);
    M2 Ao;
    Ao.a[0][0] = (Om[0][1] + d01) * dt;
    Ao.a[0][1] = Om[1][0] * dt;
    Ao.a[1][0] = Om[0][1] * dt;
    Ao.a[1][1] = (Om[1][0] + d10) * dt;
    const M2 Eo = expm2(Ao);
    const cd c0 = rho.a[0][1], c1 = rho.a[1][0];
    rho.a[0][1] = Eo.a[0][0] * c0 + Eo.a[0][1] * c1;
    rho.a[1][0] = Eo.a[1][0] * c0 + Eo.a[1][1] * c1;
}

// floquet_integrate (src/integrator.cpp:869-889): variance = the summed
// covariance spectrum (re, im pairs, NW x 3 x 3), normalised by dt/n_samples
// (CovarianceSpectrum::normalize, src/floquet.cpp:81-85); its diagonal gives
// the noise spectra of sigma_x, sigma_y, sigma_z (extract, :87-103).
inline d3 integrate(const Diag& fd, const double* variance, double n_samples, double dt, const d3& yi, double t0,
                    double tf, double w) {
    const double scale = dt / n_samples;
    cd Zeta[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Om[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    double Delta[2][2][NK];
    cd X[2][2][NK], Gamma[2][2][NK];
    M2 rho = bloch_to_density(yi, fd.f_modes_0);
    const double period = fd.dt * fd.n_prop;
    for (int i = 0; i < 3; ++i) {
        Quat c_op{0.0, i == 0 ? 1.0 : 0.0, i == 1 ? 1.0 : 0.0, i == 2 ? 1.0 : 0.0};
        cd power[NW];
        for (int j = 0; j < NW; ++j) {
            const double* v = variance + 18 * j + 6 * i + 2 * i;
            power[j] = cd(v[0], v[1]) * scale;
        }
        rates(fd.f_modes_0, fd.f_energies, c_op, fd.props.data(), fd.n_prop, period, fd.freq, power, Delta, X, Gamma,
              Zeta, Om);
    }
    markov(t0, tf, rho, Zeta, Om);
    const int n_period = (int)std::round((tf - t0) * w / (2 * M_PI));
    return density_to_bloch(rho, q_mult(fd.f_modes_0, qpow(fd.f_energies, n_period)));
}

}  // namespace floquet
}  // namespace st
