// st_spectrum.h -- the spectral path (integratorType 6) on the GPU.
//
// With integratorType 6 the reference does not integrate the spin: along each
// trajectory segment it samples the noise field gamma*(v x E/c^2 + G.r +
// a sin(w t))/2 at the multiples of the particle's step size h
// (integrateSpectrum, src/integrator.cpp:730-743; findNoiseTerm :72-80) and
// feeds every sample to NW = 17 Goertzel filters, one per Floquet sampling
// frequency, that also accumulate the imaginary diagonal of the noise
// covariance (SpectrumAggregator::update, src/floquet.cpp:111-118).  At the
// end of each interval the per-particle covariance spectra are summed into
// the run's spectrum (src/particle.cpp:824-834), which the host turns into
// the final Bloch vector (floquet_integrate, src/integrator.cpp:869-889).
//
// Device layout: the aggregator of particle i is 9*NW doubles, field f of
// it at agg[agg_index(i, f)] -- tiles of 64 particles, field-major inside a
// tile (a wave's lanes on neighbouring particles read whole lines, and a
// lane reaches every field from one base pointer at constant offsets f*512
// bytes): s1 of frequency q and axis c at f = 3q + c, s2 at 3NW + 3q + c,
// imag_diag at 6NW + 3q + c; its sample count (the reference's int) is
// nsamp[i].
//
// The sampling kernel reads the pipeline's segment records like the spin
// kernel.  A lane holds FB filters in registers (9*FB doubles) and the
// frequencies are split over ceil(NW/FB) launches, one per register block,
// each walking every record of the chunk again: the filter constants are
// then wave-uniform kernel arguments (scalar loads) and a launch's block
// (6 filters, 108 VGPRs of state) fits the register file without spills at
// two waves per SIMD; the 18 independent filter chains per lane carry the
// FP64 latency.  In the exact variant each sample costs the noise term (55
// FP64 operations, three sines when the test noise is on) plus 11 operations
// per filter and axis, in the reference's operation order; the fast variant
// (below) 7 and 4.
#pragma once

#include "st_pipeline.h"

namespace st {

constexpr int NW = 17;        // include/floquet.h:34
constexpr int NK = 11;        // include/floquet.h:33
constexpr int AGG_FIELDS = 9 * NW;
ST_HD __forceinline__ uint64_t agg_tile(uint32_t pid) { return (uint64_t)(pid >> 6) * (AGG_FIELDS * 64) + (pid & 63u); }
ST_HD __forceinline__ uint64_t agg_index(uint32_t pid, int f) { return agg_tile(pid) + (uint64_t)f * 64; }
constexpr int NW_PAD = 18;    // filter constants padded with zeros to whole register blocks
constexpr int SPEC_FB = 6;    // filters per register block
constexpr int SPEC_NB = (NW + SPEC_FB - 1) / SPEC_FB;
// samples per loop trip: eight for the fast variant; the exact variant's
// longer sample body spills registers beyond four
template <int VAR>
constexpr int spec_unroll() { return VAR == VAR_FAST ? 8 : 4; }
constexpr int SPEC_ENTRIES = NW * 18;   // one interval's spectrum: NW 3x3 complex matrices
static_assert(4 + SPEC_NB <= 8, "the spectral blocks' queue heads sit in d_small[4..7]");
static_assert(SPEC_NB * SPEC_FB <= NW_PAD, "padded filter constants cover every block");

struct SpecArgs {
    double* agg;
    uint32_t* nsamp;
    double k2[NW_PAD], cw[NW_PAD], sw[NW_PAD];   // 2cos(w dt), cos(w dt), sin(w dt) of each filter (host libm)
    double sw2[NW_PAD];          // 2 sin(w dt) (fast variant)
    double na[3], nw[3];         // test-noise amplitudes and angular frequencies
    int32_t noise_on;            // some amplitude is non-zero
    int32_t block;               // this launch's register block
};

// findNoiseTerm (src/integrator.cpp:72-80): interpolate -> grad ->
// (1/c^2) v x E + G.r + a sin(w t) -> gamma*B, in the reference's order.
// Without test noise the sine term is +0 (the reference adds 0*sin(0*t)).
__device__ __forceinline__ d3 noise_exact(const SimParams& P, const Phys& F, const SpecArgs& S, double t,
                                          const Seg& G) {
    const double d = t - G.t0;
    d3 p, v;
    if (P.gravity) {
        p.x = (G.p0.x + G.v0.x * d) + 0.0;
        p.y = (G.p0.y + G.v0.y * d) + (HALF_G * d) * d;
        p.z = (G.p0.z + G.v0.z * d) + 0.0;
        v.x = G.v0.x + 0.0;
        v.y = G.v0.y + G_CONST * d;
        v.z = G.v0.z + 0.0;
    } else {
        p.x = G.p0.x + G.v0.x * d;
        p.y = G.p0.y + G.v0.y * d;
        p.z = G.p0.z + G.v0.z * d;
        v = G.v0;
    }
    const double r[3] = {p.x, p.y, p.z};
    double gg[3];
    st_field_grad(r, F.G, gg);   // the grad hook of spintrack_fields.h
    const double gx = gg[0], gy = gg[1], gz = gg[2];
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (S.noise_on) {
        nx = S.na[0] * sin(S.nw[0] * t);
        ny = S.na[1] * sin(S.nw[1] * t);
        nz = S.na[2] * sin(S.nw[2] * t);
    }
    const double cx = v.y * F.Ez - v.z * F.Ey;
    const double cy = v.z * F.Ex - v.x * F.Ez;
    const double cz = v.x * F.Ey - v.y * F.Ex;
    const double Bx = (cx * IC2 + gx) + nx;
    const double By = (cy * IC2 + gy) + ny;
    const double Bz = (cz * IC2 + gz) + nz;
    return d3{Bx * F.gamma, By * F.gamma, Bz * F.gamma};
}

// Fast variant: along a segment the halved noise field without test noise
// is a quadratic in tau = t - t0 -- position is quadratic and velocity
// linear in tau, and every term is linear in them:
//   x(tau) = gamma/2 [ (v0 x E)/c^2 + G p0 + tau ((g y^ x E)/c^2 + G v0)
//                      + tau^2 (g/2) G y^ ],  y^ x E = (Ez, 0, -Ex).
struct NoisePoly {
    d3 w0, w1, w2;
};
__device__ __forceinline__ NoisePoly noise_poly(const SimParams& P, const Phys& F, const Seg& G) {
    const double hg = 0.5 * F.gamma;
    const double g = P.gravity ? G_CONST : 0.0;
    const d3 p = G.p0, v = G.v0;
    const double cx = v.y * F.Ez - v.z * F.Ey, cy = v.z * F.Ex - v.x * F.Ez, cz = v.x * F.Ey - v.y * F.Ex;
    NoisePoly q;
    q.w0 = d3{hg * (cx * IC2 + (p.x * F.G[0] + p.y * F.G[1] + p.z * F.G[2])),
              hg * (cy * IC2 + (p.x * F.G[3] + p.y * F.G[4] + p.z * F.G[5])),
              hg * (cz * IC2 + (p.x * F.G[6] + p.y * F.G[7] + p.z * F.G[8]))};
    q.w1 = d3{hg * (g * F.Ez * IC2 + (v.x * F.G[0] + v.y * F.G[1] + v.z * F.G[2])),
              hg * (v.x * F.G[3] + v.y * F.G[4] + v.z * F.G[5]),
              hg * (-g * F.Ex * IC2 + (v.x * F.G[6] + v.y * F.G[7] + v.z * F.G[8]))};
    q.w2 = d3{hg * (0.5 * g) * F.G[1], hg * (0.5 * g) * F.G[4], hg * (0.5 * g) * F.G[7]};
    return q;
}

// first_sample_point (src/integrator.cpp:724-728)
__device__ __forceinline__ double first_sample_point(double t0, double h) { return (floor(t0 / h) + 1) * h; }

// One launch per register block: every lane walks its particle's records of
// this chunk, sampling the noise field at the multiples of tempH inside each
// integrated segment into the block's filters.  Block 0 counts the samples;
// the last block also does the spin side's per-record bookkeeping (the spin
// is not integrated: S only turns NaN at a stop and is written at interval
// ends), so t_spin and tempH are read-only for the other blocks.
//
// The record after the one being sampled is loaded when sampling starts, so
// its latency hides behind the samples instead of stalling the wave at the
// segment change.  Exact variant: the reference's operations in its order.
// Fast variant: the noise from its segment polynomial (two FMAs per axis)
// and per filter and axis s1' = fma(2cos, s1, x) - s2, imag += (2 sin s2') x
// as one FMA -- the reference's "- re*0" term only matters for non-finite
// filter states -- 4 operations instead of 11; agrees with the exact
// variant to rounding.
template <int FB, int VAR>
__global__ __launch_bounds__(256, 2) void k_spectrum(const SimParams P, const DevState D, const PipeArgs A,
                                                     const SpecArgs S) {
    const RunArgs& R = A.R;
    const Phys& F = A.phys[0];
    const int q0 = S.block * FB;
    const int nf = NW - q0 < FB ? NW - q0 : FB;
    const bool last = S.block == SPEC_NB - 1;
    const bool first = S.block == 0;
    SpinLane L;
    double s1[FB][3], s2[FB][3], ac[FB][3];
    Seg seg;
    NoisePoly poly;
    SegRec nxt;            // the record after L.j, loaded ahead
    double ts = 0.0, lim = 0.0;
    // first_sample_point(t_key, h_key): a segment's end value is the next
    // segment's start value (one division per record instead of two)
    double t_key = __builtin_nan(""), h_key = 0.0, fsp_key = 0.0;
    uint32_t ns = 0;
    unsigned long long cnt[4] = {0, 0, 0, 0};   // samples, -, segments, trips
    enum { S_DONE = 0, S_NEXT = 1, S_SAMPLE = 2 };
    int phase = S_DONE;
    const uint32_t n_list = list_len(A);
    // the block's filter constants, read once (a padding filter has zeros)
    double ck[FB], cc[FB], cs[FB];
#pragma unroll
    for (int f = 0; f < FB; ++f) {
        ck[f] = S.k2[q0 + f];
        cc[f] = VAR == VAR_FAST ? S.sw2[q0 + f] : S.cw[q0 + f];
        cs[f] = S.sw[q0 + f];
    }
    auto load = [&](uint32_t pid) {
        spin_load<false>(D, A, pid, L);
        ns = 0;
        const double* a = S.agg + agg_tile(pid) + (uint64_t)(3 * q0) * 64;
#pragma unroll
        for (int f = 0; f < FB; ++f)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool in = f < nf;
                s1[f][c] = in ? a[(3 * f + c) * 64] : 0.0;
                s2[f][c] = in ? a[(3 * NW + 3 * f + c) * 64] : 0.0;
                ac[f][c] = in ? a[(6 * NW + 3 * f + c) * 64] : 0.0;
            }
        // the particle's state has arrived before the record prefetch is
        // issued: the sample loop then never waits on anything but the
        // prefetched record, at the next segment change
        __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0)
        if (L.nrec > 0) nxt = get_rec(A.recs, A.C, pid, 0);
    };
    auto store = [&]() {
        const uint32_t pid = L.pid;
        double* a = S.agg + agg_tile(pid) + (uint64_t)(3 * q0) * 64;
#pragma unroll
        for (int f = 0; f < FB; ++f)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (f < nf) {
                    a[(3 * f + c) * 64] = s1[f][c];
                    a[(3 * NW + 3 * f + c) * 64] = s2[f][c];
                    a[(6 * NW + 3 * f + c) * 64] = ac[f][c];
                }
            }
        if (first) {
            S.nsamp[pid] += ns;
            D.st<uint64_t>(F_ACC, pid * 8u, D.ld<uint64_t>(F_ACC, pid * 8u) + ns);
        }
        if (last) {
            const uint32_t o = pid * 8u;
            D.st<uint64_t>(F_SSTATE, o, (uint64_t)(uint32_t)L.k | ((uint64_t)L.stopped << 33));
            D.st(F_TEMPH, o, L.tempH);
            D.st(F_TSPIN, o, L.t_spin);
            D.st(F_SX, o, L.S.x); D.st(F_SY, o, L.S.y); D.st(F_SZ, o, L.S.z);
            D.st<uint64_t>(F_SEG, o, D.ld<uint64_t>(F_SEG, o) + L.seg);
        }
    };
    // SpectrumAggregator::update(B/2) at time t: goertzel_stage_1, then
    // goertzel_stage_2_vector of the new state, then covMat::add_outer(a, b,
    // x, 0) on imag_diag
    // One sample with the Goertzel state held as (sa = s1, sb = s2): the new
    // s1 is written over sb, after which sb holds s1 and sa holds s2.  The
    // trip loop calls it with the two arrays' roles alternating, so the state
    // rotation s2 = s1, s1 = new costs no register copies (under a lane mask
    // each copy is a real move: four 32-bit moves per filter and axis).
    auto sample = [&](double t, double (&sa)[FB][3], double (&sb)[FB][3]) {
        double x[3];
        if constexpr (VAR == VAR_FAST) {
            const double tau = t - seg.t0;
            x[0] = fma(tau, fma(tau, poly.w2.x, poly.w1.x), poly.w0.x);
            x[1] = fma(tau, fma(tau, poly.w2.y, poly.w1.y), poly.w0.y);
            x[2] = fma(tau, fma(tau, poly.w2.z, poly.w1.z), poly.w0.z);
            if (S.noise_on) {
                const double hg = 0.5 * F.gamma;
                x[0] = fma(hg * S.na[0], sin(S.nw[0] * t), x[0]);
                x[1] = fma(hg * S.na[1], sin(S.nw[1] * t), x[1]);
                x[2] = fma(hg * S.na[2], sin(S.nw[2] * t), x[2]);
            }
        } else {
            const d3 B = noise_exact(P, F, S, t, seg);
            x[0] = B.x / 2;
            x[1] = B.y / 2;
            x[2] = B.z / 2;
        }
#pragma unroll
        for (int f = 0; f < FB; ++f) {
            if constexpr (VAR == VAR_FAST) {
                const double k = ck[f], sw2 = cc[f];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    sb[f][c] = fma(k, sa[f][c], x[c]) - sb[f][c];
                    ac[f][c] = fma(sw2 * sa[f][c], x[c], ac[f][c]);
                }
            } else {
                const double k = ck[f], cw = cc[f], sw = cs[f];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    sb[f][c] = (x[c] + k * sa[f][c]) - sb[f][c];
                    const double re = sb[f][c] - cw * sa[f][c];
                    const double im = sw * sa[f][c];
                    ac[f][c] = ac[f][c] + 2 * (im * x[c] - re * 0.0);
                }
            }
        }
    };
    {
        const uint32_t q = acquire_index(R.queue);
        if (q < n_list) {
            load(A.list[q]);
            phase = S_NEXT;
        }
    }
    while (phase != S_DONE) {
        ++cnt[3];
        if (phase == S_NEXT) {
            while (phase == S_NEXT) {
                if (L.j >= L.nrec) {
                    cnt[0] += ns; cnt[2] += L.seg;
                    store();
                    phase = S_DONE;
                    const uint32_t q = acquire_index(R.queue);
                    if (q < n_list) {
                        load(A.list[q]);
                        phase = S_NEXT;
                        continue;
                    }
                    break;
                }
                const SegRec rec = nxt;
                if (L.j + 1 < L.nrec) nxt = get_rec(A.recs, A.C, L.pid, L.j + 1);
                if (rec.flags & R_FIRST) L.tempH = L.h;
                if ((rec.flags & R_INTEGRATE) && !L.stopped) {
                    seg.t0 = L.t_spin;
                    seg.p0 = d3{rec.px, rec.py, rec.pz};
                    seg.v0 = d3{rec.vx, rec.vy, rec.vz};
                    if (VAR == VAR_FAST) poly = noise_poly(P, F, seg);
                    ts = (L.t_spin == t_key && L.tempH == h_key) ? fsp_key : first_sample_point(L.t_spin, L.tempH);
                    fsp_key = first_sample_point(rec.t, L.tempH);
                    t_key = rec.t;
                    h_key = L.tempH;
                    lim = fsp_key - (L.tempH / 2);
                    L.seg++;
                    if (ts < lim) phase = S_SAMPLE;
                }
                if (last) {
                    if (rec.flags & R_STOP) {
                        const double nan = __builtin_nan("");
                        L.S = d3{nan, nan, nan};
                    }
                    if (rec.flags & R_IEND) {
                        spin_block(R, L);
                        L.k++;
                    }
                }
                if (!(rec.flags & R_EMPTY)) L.t_spin = rec.t;
                L.j++;
            }
        }
        if (phase == S_SAMPLE) {
            // up to spec_unroll samples per trip: the divergent record
            // change above is paid once per several samples, not per sample
            // Samples alternate the roles of s1 and s2 (even: s1 = sa, odd:
            // s1 = sb).  A lane that stops sampling in this trip stays stopped
            // until its end, so after an odd number of samples its two arrays
            // are swapped back once here.
            const uint32_t ns0 = ns;
            auto advance = [&]() {
                ++ns;
                const double tn = ts + L.tempH;
                // (a step below the rounding of ts would never end the
                // segment; the host bounds h so that this cannot happen)
                if (tn == ts) lim = ts;
                ts = tn;
            };
            static_assert(spec_unroll<VAR>() % 2 == 0, "samples are taken in pairs of roles");
#pragma unroll
            for (int u = 0; u < spec_unroll<VAR>(); u += 2) {
                if (u == 0 || ts < lim) {
                    sample(ts, s1, s2);
                    advance();
                }
                if (ts < lim) {
                    sample(ts, s2, s1);
                    advance();
                }
            }
            if ((ns - ns0) & 1u) {
#pragma unroll
                for (int f = 0; f < FB; ++f)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double a = s1[f][c];
                        s1[f][c] = s2[f][c];
                        s2[f][c] = a;
                    }
            }
            if (!(ts < lim)) phase = S_NEXT;
        }
    }
    if (first) flush_counters(R.work, cnt);
}

// ---- interval sums ----------------------------------------------------------
// get_covariance_spectrum (src/floquet.cpp:131-147) of particle pid,
// frequency q: out[6*r + 2*s + {0: re, 1: im}] of the 3x3 matrix
// c c^H - i diag(imag_diag), c = (s1 - cos s2) + i sin s2.
__device__ __forceinline__ void cov_entries(const SpecArgs& S, uint32_t pid, int q, double (&out)[18]) {
    const double cw = S.cw[q], sw = S.sw[q];
    double re[3], im[3], id[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double a = S.agg[agg_index(pid, 3 * q + c)];
        const double b = S.agg[agg_index(pid, 3 * NW + 3 * q + c)];
        id[c] = S.agg[agg_index(pid, 6 * NW + 3 * q + c)];
        re[c] = a - cw * b;
        im[c] = sw * b;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            double vr = re[r] * re[s] - im[r] * (-im[s]);
            double vi = re[r] * (-im[s]) + im[r] * re[s];
            if (r == s) {
                vr = vr + 0.0;
                vi = vi + (-id[r]);
            }
            out[6 * r + 2 * s] = vr;
            out[6 * r + 2 * s + 1] = vi;
        }
}

// The reference CPU path's sum (src/particle.cpp:824-834): its in-place tree
// leaves particles 0, 1, 2, 4, ... < n in slot 0, added in that order, and
// slot 0 is then added to the run's spectrum.  Thread q replays that order
// for frequency q; thread NW sums the sample counts.
__global__ __launch_bounds__(64) void k_spec_partial(const SpecArgs S, uint32_t n, double* acc,
                                                    unsigned long long* acc_n) {
    const int q = threadIdx.x;
    if (q < NW) {
        double t[18], c[18];
        cov_entries(S, 0, q, t);
        for (uint32_t s = 1; s < n; s *= 2) {
            cov_entries(S, s, q, c);
#pragma unroll
            for (int e = 0; e < 18; ++e) t[e] = t[e] + c[e];
        }
#pragma unroll
        for (int e = 0; e < 18; ++e) acc[18 * q + e] = acc[18 * q + e] + t[e];
    } else if (q == NW) {
        unsigned long long t = S.nsamp[0];
        for (uint32_t s = 1; s < n; s *= 2) t += S.nsamp[s];
        *acc_n += t;
    }
}

// Every particle (aggregateSpectrum, src/particle.cpp:893-897), as a
// deterministic two-level sum: workgroup g adds its particles' entries in a
// fixed tree and writes part[g][SPEC_ENTRIES + 1]; k_spec_final adds the
// partials in workgroup order.  (The reference adds particle by particle;
// the sums agree to rounding.)
constexpr int SPEC_RED_BLOCKS = 512;
__global__ __launch_bounds__(256) void k_spec_all(const SpecArgs S, uint32_t n, double* part) {
    __shared__ double red[18][4];
    const uint32_t gstride = gridDim.x * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* mine = part + (uint64_t)blockIdx.x * (SPEC_ENTRIES + 1);
    for (int q = 0; q < NW; ++q) {
        double v[18], c[18];
#pragma unroll
        for (int e = 0; e < 18; ++e) v[e] = 0.0;
        for (uint32_t pid = tid; pid < n; pid += gstride) {
            cov_entries(S, pid, q, c);
#pragma unroll
            for (int e = 0; e < 18; ++e) v[e] += c[e];
        }
#pragma unroll
        for (int e = 0; e < 18; ++e) {
            double w = v[e];
            for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
            if (lane == 0) red[e][wave] = w;
        }
        __syncthreads();
        if (threadIdx.x < 18) {
            const int e = threadIdx.x;
            mine[18 * q + e] = ((red[e][0] + red[e][1]) + red[e][2]) + red[e][3];
        }
        __syncthreads();
    }
    double m = 0.0;
    for (uint32_t pid = tid; pid < n; pid += gstride) m += (double)S.nsamp[pid];
    for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off);
    if (lane == 0) red[0][wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) mine[SPEC_ENTRIES] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
}

__global__ __launch_bounds__(320) void k_spec_final(const double* part, int nblocks, double* acc,
                                                   unsigned long long* acc_n) {
    const int e = threadIdx.x;
    if (e > SPEC_ENTRIES) return;
    double t = 0.0;
    for (int g = 0; g < nblocks; ++g) t += part[(uint64_t)g * (SPEC_ENTRIES + 1) + e];
    if (e < SPEC_ENTRIES) acc[e] = acc[e] + t;
    else *acc_n += (unsigned long long)t;
}

// Clears the aggregators and sample counts of particles 0..n-1 (the reset at
// the start of each interval, src/particle.cpp:746-747); a fill kernel rather
// than a memset node, so the clear is safe inside captured work.
__global__ __launch_bounds__(256) void k_spec_clear(const SpecArgs S, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int f = 0; f < AGG_FIELDS; ++f) S.agg[agg_index(i, f)] = 0.0;
    S.nsamp[i] = 0;
}

}  // namespace st
