// st_adjoint.h -- adjoint (reverse-mode) Bloch integration along recorded
// trajectories.
//
// The forward spin obeys S' = S x omega(t), omega = gamma*(B0 + G.r +
// (1/c^2) v x E + pulse).  For a loss L that depends on the spins at the
// output times through the ensemble means, L = l(<S>(t_1), ..., <S>(t_K)),
// the adjoint lambda of one particle obeys the same equation backwards,
// lambda' = lambda x omega, and jumps by g_k = dl/d<S>(t_k) at every output
// time.  Both lambda and S are transported by the same rotation, so
// u = lambda x S obeys u' = u x omega as well, and
//
//   dL/dtheta = (1/N) sum_particles int u . (d omega / d theta) dt,
//
// with jumps u(t_k-) = u(t_k+) + g_k x S(t_k).  Only u (3 components) is
// integrated, backwards over every segment, with the DOP853 controller; the
// gradient integrals ride along as a quadrature.  omega is linear in
// theta = (G, E, B0):
//   dL/dG_ij = gamma int u_i r_j dt,  dL/dE = gamma/c^2 int u x v dt,
//   dL/dB0 = gamma int u dt,
// and along one segment r and v are polynomials in tau = t - t0, so the
// kernel only accumulates the moments int u tau^n dt (n = 0, 1, 2) per
// segment and folds them with the segment's p0, v0 and gravity.
//
// The trajectory is recomputed, not stored: the forward pipeline keeps the
// kinematic state at every chunk start (every C segments of each particle);
// the backward pass replays one chunk's kinematics (k_kin, bit-identical
// records) and integrates that chunk's records in reverse.
#pragma once

#include "st_pipeline.h"

namespace st {

// per-particle adjoint state, one 8-byte array per field (same layout rule
// as DevState)
enum AdjField : int {
    A_UX, A_UY, A_UZ,   // u = lambda x S at the current backward time
    A_H,                // carried step size
    A_K,                // interval index of the next output time going backwards
    A_FAIL,             // nonzero: a backward segment failed (integrator status)
    A_G0,               // 15 accumulators: int u_i r_j dt (9, row-major i, j),
                        // int u x v dt (3), int u dt (3)
    A_COUNT = A_G0 + 15
};
constexpr int N_GRAD = 15;

struct AdjArgs {
    const char* recs;        // records of the replayed chunk
    uint32_t C;
    const uint32_t* list;    // particles of the chunk
    uint32_t n_list;
    const uint32_t* seg_count;
    DevState ck;             // kinematic state at the chunk start (checkpoint)
    DevState adj;            // adjoint state
    const char* blocks;      // forward snapshot blocks (S at the output times)
    uint64_t block_bytes;
    uint32_t n;              // particles per block
    const double* seeds;     // g_k, 3 per output time (per configuration for batches)
    uint32_t n_targets;
    const Phys* phys;        // batched handles
};

__global__ __launch_bounds__(256) void k_adj_begin(const DevState A, uint32_t n, int32_t k_last, double h0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = i * 8u;
    A.st(A_UX, o, 0.0); A.st(A_UY, o, 0.0); A.st(A_UZ, o, 0.0);
    A.st(A_H, o, h0);
    A.st<int64_t>(A_K, o, k_last);
    A.st<int64_t>(A_FAIL, o, 0);
    for (int f = 0; f < N_GRAD; ++f) A.st(A_G0 + f, o, 0.0);
}

// PULSE = false: the fast trial's pulse-free form (runs without a pulse term)
template <bool BATCH, bool PULSE>
__global__ __launch_bounds__(256) void k_adj(const SimParams P, const DevState D, const AdjArgs A) {
#pragma clang fp contract(fast)
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= A.n_list) return;
    const uint32_t pid = A.list[li];
    const uint32_t o = pid * 8u;
    const Phys& F = lane_phys<BATCH>(P, D, A.phys, pid);
    d3 u{A.adj.ld(A_UX, o), A.adj.ld(A_UY, o), A.adj.ld(A_UZ, o)};
    double h = A.adj.ld(A_H, o);
    int64_t kb = A.adj.ld<int64_t>(A_K, o);
    int64_t fail = A.adj.ld<int64_t>(A_FAIL, o);
    const double ay = P.gravity ? G_CONST : 0.0;
    const uint32_t nrec = A.seg_count[pid];
    const double t_chunk0 = A.ck.ld(F_T, o);
    uint32_t nacc = 0, nrej = 0;
    for (int32_t j = (int32_t)nrec - 1; j >= 0; --j) {
        const SegRec rec = get_rec(A.recs, A.C, pid, (uint32_t)j);
        if ((rec.flags & R_IEND) && kb >= 0) {
            // the output time t_k closes this record: u += g_k x S(t_k)
            const double* Sk = reinterpret_cast<const double*>(A.blocks + (uint64_t)kb * A.block_bytes + 56ull * A.n) +
                               3ull * pid;
            const uint64_t cfg = BATCH ? D.ld<uint64_t>(F_CFG, o) : 0;
            const double* gp = A.seeds + 3 * (cfg * A.n_targets + (uint64_t)kb);
            const d3 gk{gp[0], gp[1], gp[2]};
            u = d3{u.x + (gk.y * Sk[2] - gk.z * Sk[1]), u.y + (gk.z * Sk[0] - gk.x * Sk[2]),
                   u.z + (gk.x * Sk[1] - gk.y * Sk[0])};
            --kb;
            // a non-finite spin (a particle that failed forward) cannot be
            // integrated: flag it instead of running the controller on NaN
            if (!(__builtin_isfinite(u.x) && __builtin_isfinite(u.y) && __builtin_isfinite(u.z))) fail = -3;
        }
        if (!(rec.flags & R_INTEGRATE) || fail) continue;
        const double t0 = (j > 0) ? get_rec(A.recs, A.C, pid, (uint32_t)(j - 1)).t : t_chunk0;
        const double Ts = rec.t - t0;
        if (!(Ts > 0.0)) continue;
        Seg seg;
        seg.t0 = t0;
        seg.p0 = d3{rec.px, rec.py, rec.pz};
        seg.v0 = d3{rec.vx, rec.vy, rec.vz};
        // omega along the segment in tau = t - t0, then the backward field
        // -omega(t_end - s) as a polynomial in s
        const Poly q = make_poly(P, F, seg);
        Poly qb;
        qb.w0 = d3{-(q.w0.x + Ts * (q.w1.x + Ts * q.w2.x)), -(q.w0.y + Ts * (q.w1.y + Ts * q.w2.y)),
                   -(q.w0.z + Ts * (q.w1.z + Ts * q.w2.z))};
        qb.w1 = d3{q.w1.x + 2.0 * Ts * q.w2.x, q.w1.y + 2.0 * Ts * q.w2.y, q.w1.z + 2.0 * Ts * q.w2.z};
        qb.w2 = d3{-q.w2.x, -q.w2.y, -q.w2.z};
        qb.pag = q.pag;
        qb.pw = q.pw;
        qb.paxis = q.paxis;
        qb.pon = q.pon;
        const Seg sb{0.0, seg.p0, seg.v0};
        AdjQuad Q;
        Q.n0 = Q.n1 = Q.n2 = d3{0.0, 0.0, 0.0};
        Q.t_end = rec.t;
        DopState dop;
        dop.x = 0.0;
        dop.xf = Ts;
        dop.facold = 1.0E-4;
        dop.posneg = 1.0;
        dop.last = 0;
        dop.reject = 0;
        dop.nstep = 0;
        dop.k1 = poly_rhs(0.0, u, qb, qb.pon ? -(qb.pag * cos(qb.pw * rec.t)) : 0.0);
        int st = 0;
        while ((st = dop_trial_fast_t<true, PULSE>(P, F, dop, u, h, sb, qb, nacc, nrej, &Q)) == 0) {
        }
        if (st < 0) {
            fail = st;
            continue;
        }
        // s-moments -> tau-moments (tau = Ts - s)
        const d3 M0 = Q.n0;
        const d3 M1{Ts * Q.n0.x - Q.n1.x, Ts * Q.n0.y - Q.n1.y, Ts * Q.n0.z - Q.n1.z};
        const d3 M2{Ts * (Ts * Q.n0.x - 2.0 * Q.n1.x) + Q.n2.x, Ts * (Ts * Q.n0.y - 2.0 * Q.n1.y) + Q.n2.y,
                    Ts * (Ts * Q.n0.z - 2.0 * Q.n1.z) + Q.n2.z};
        // the segment's contributions (accumulated in HBM: a segment spans
        // tens of steps, and keeping 15 sums live in registers would halve
        // the occupancy)
        double dg[N_GRAD];
        // int u_i r_j dt with r = p0 + v0 tau + (1/2) a tau^2, a = (0, ay, 0)
        const double Mi[3][3] = {{M0.x, M1.x, M2.x}, {M0.y, M1.y, M2.y}, {M0.z, M1.z, M2.z}};
        const double p[3] = {seg.p0.x, seg.p0.y, seg.p0.z}, v[3] = {seg.v0.x, seg.v0.y, seg.v0.z};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const double a_j = (jj == 1) ? 0.5 * ay : 0.0;
                dg[3 * i + jj] = p[jj] * Mi[i][0] + v[jj] * Mi[i][1] + a_j * Mi[i][2];
            }
        }
        // int u x v dt with v = v0 + a tau
        dg[9] = (M0.y * seg.v0.z - M0.z * seg.v0.y) + (-M1.z * ay);
        dg[10] = (M0.z * seg.v0.x - M0.x * seg.v0.z);
        dg[11] = (M0.x * seg.v0.y - M0.y * seg.v0.x) + (M1.x * ay);
        dg[12] = M0.x;
        dg[13] = M0.y;
        dg[14] = M0.z;
#pragma unroll
        for (int f = 0; f < N_GRAD; ++f) A.adj.st(A_G0 + f, o, A.adj.ld(A_G0 + f, o) + dg[f]);
    }
    A.adj.st(A_UX, o, u.x); A.adj.st(A_UY, o, u.y); A.adj.st(A_UZ, o, u.z);
    A.adj.st(A_H, o, h);
    A.adj.st<int64_t>(A_K, o, kb);
    A.adj.st<int64_t>(A_FAIL, o, fail);
}

// Per-block partial sums of the 15 accumulators (deterministic order; the
// host adds the partials in block order).  Batched handles sum per
// configuration: partial[(block * n_cfg + cfg) * 15 + f] -- blocks never
// straddle configurations because the grid is laid out per configuration.
__global__ __launch_bounds__(256) void k_adj_partials(const DevState A, const uint32_t* cfg_off, uint32_t blocks_per_cfg,
                                                      double* partial) {
    const uint32_t cfg = blockIdx.y;
    const uint32_t lo = cfg_off[cfg], hi = cfg_off[cfg + 1];
    __shared__ double red[256];
    double acc[N_GRAD];
#pragma unroll
    for (int f = 0; f < N_GRAD; ++f) acc[f] = 0.0;
    for (uint32_t i = lo + blockIdx.x * 256 + threadIdx.x; i < hi; i += blocks_per_cfg * 256) {
#pragma unroll
        for (int f = 0; f < N_GRAD; ++f) acc[f] += A.ld(A_G0 + f, i * 8u);
    }
    for (int f = 0; f < N_GRAD; ++f) {
        red[threadIdx.x] = acc[f];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[((uint64_t)cfg * blocks_per_cfg + blockIdx.x) * N_GRAD + f] = red[0];
        __syncthreads();
    }
}

}  // namespace st
