// st_engine.hip -- gfx950 kernels of the spin-tracking hot path and the C ABI
// declared in include/spintrack.h.
//
// Design (one particle per lane, wave64):
//  * Particle state lives in HBM as structure-of-arrays (one array per scalar
//    component), so the lanes of a wave load and store it coalesced.
//  * The lane engine flattens the reference's three nested loops (output
//    interval -> event -> DOP853/RK45 trial step; src/simulation.cpp:84-92,
//    src/particle.cpp:754-803, src/integrator.cpp:128-247) into ONE loop whose
//    every trip performs at most one event (cheap, branchy) and exactly one
//    integrator trial step (the expensive, uniform FP64 work).  Lanes in
//    different segments, intervals or particles therefore stay convergent on
//    the trial step, which is where the time goes.
//  * The plain kernel (st_advance_to) maps lane i to particle i for one
//    output interval.  The persistent kernel (st_run) keeps a grid that just
//    fills the chip, carries each particle through all output times without a
//    grid barrier, writes snapshot blocks as each lane crosses an output time,
//    and refills a lane from an atomic work queue when its particle is done.
//  * Tableau coefficients are compile-time constants (SGPR literals); run
//    parameters are one uniform kernel-argument block.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/spintrack.h"
#include "st_device.h"
#include "st_integrators.h"

static_assert(sizeof(st_options) == 440, "st_options must match the reference's 440-byte options");
static_assert(offsetof(st_options, nmax) == 360, "st_options layout");
static_assert(offsetof(st_options, dist) == 384, "st_options layout");
static_assert(offsetof(st_options, noiseAmplitudes) == 392, "st_options layout");
static_assert(sizeof(st_rng) == 48, "st_rng must match the reference's 48-byte rngState");

namespace st {

// ---------------------------------------------------------------------------
// device-resident state: one allocation, one 8-byte array per field, field f
// of particle i at base + f*stride + 8*i.  Passing one base pointer instead
// of ~40 array pointers keeps the kernel arguments small enough for SGPRs.
// ---------------------------------------------------------------------------
enum Field : int {
    F_T, F_TOLD, F_NEXTGAS, F_H, F_TF,
    F_PX, F_PY, F_PZ, F_POX, F_POY, F_POZ,
    F_VX, F_VY, F_VZ, F_VOX, F_VOY, F_VOZ,
    F_SX, F_SY, F_SZ,
    F_RX, F_RY, F_RZ, F_RW, F_SPARE,
    F_FLAGS,      // failure (low 32 bits) | stop<<32 | hasSpare<<33 | coll_type<<40 | wall_hit<<48
    F_NBOUNCE, F_NCOLL, F_NSTEPS, F_PARTID,
    F_ACC, F_REJ, F_SEG,   // per-particle ODE counters (accepted, rejected trials, segments)
    F_COUNT
};

struct DevState {
    char* base;
    uint64_t stride;   // bytes between consecutive fields
    template <typename T = double>
    __device__ __host__ __forceinline__ T* f(int field) const {
        return reinterpret_cast<T*>(base + (uint64_t)field * stride);
    }
    // Lane access through a uniform field base plus a 32-bit byte offset
    // (global_load/store saddr form).  Indexing with a 64-bit per-lane
    // address instead makes the compiler keep one address pair per field
    // alive from the load at the top of the loop to the store at its end.
    template <typename T = double>
    __device__ __forceinline__ T ld(int field, uint32_t off) const {
        return *reinterpret_cast<const T*>(base + (uint64_t)field * stride + off);
    }
    template <typename T = double>
    __device__ __forceinline__ void st(int field, uint32_t off, T v) const {
        *reinterpret_cast<T*>(base + (uint64_t)field * stride + off) = v;
    }
};

__device__ __host__ __forceinline__ uint64_t pack_flags(int32_t failure, bool stop, bool has_spare, char ct, char wh) {
    return (uint64_t)(uint32_t)failure | ((uint64_t)(stop ? 1 : 0) << 32) | ((uint64_t)(has_spare ? 1 : 0) << 33) |
           ((uint64_t)(uint8_t)ct << 40) | ((uint64_t)(uint8_t)wh << 48);
}

struct RunArgs {
    const double* targets;   // device array of output times, or nullptr
    double t_single;         // the only target when targets == nullptr
    double t_prev;           // the target before targets[0]
    const double* stop_at;   // per-particle end times or nullptr
    char* blocks;            // snapshot blocks or nullptr
    uint64_t block_bytes;
    unsigned long long* work;  // 4 counters: accepted, rejected, segments, trips
    uint32_t* queue;         // work-queue head (persistent kernel)
    uint32_t n;              // particles
    int32_t n_targets;
    uint32_t max_events;     // per-interval event cap (watchdog against runaway lanes)
    // segment trace of one particle (trace_pid), or trace == nullptr
    double* trace;
    uint32_t* trace_count;
    uint32_t trace_pid;
    uint32_t trace_cap;
};

// Trace records (TRACE_DOUBLES doubles each), two per integrated segment:
// tag (1 = segment start, 2 = segment end), t_old, t, pos_old[3], v_old[3],
// S[3], step size, accepted, rejected (this particle, this launch),
// coll_type, wall_hit, integrator status, n_bounce, n_coll.
constexpr int TRACE_DOUBLES = 20;

__device__ __noinline__ void trace_record(const RunArgs& R, int tag, double t_old, double t, d3 po, d3 vo,
                                          d3 s, double h, uint32_t acc, uint32_t rej, char ct, char wh,
                                          int status, uint64_t nb, uint64_t nc) {
    const uint32_t k = atomicAdd(R.trace_count, 1u);
    if (k >= R.trace_cap) return;
    double* r = R.trace + (uint64_t)k * TRACE_DOUBLES;
    r[0] = tag; r[1] = t_old; r[2] = t;
    r[3] = po.x; r[4] = po.y; r[5] = po.z;
    r[6] = vo.x; r[7] = vo.y; r[8] = vo.z;
    r[9] = s.x; r[10] = s.y; r[11] = s.z;
    r[12] = h; r[13] = acc; r[14] = rej; r[15] = (double)ct; r[16] = (double)wh;
    r[17] = status; r[18] = (double)nb; r[19] = (double)nc;
}

constexpr int32_t FAIL_WATCHDOG = -9;

struct Lane {
    uint32_t pid;
    int32_t k;          // index of the open output interval
    double tf, h, tempH;
    d3 S;
    Kin K;
    Rng rng;
    uint64_t n_steps;
    uint32_t n_events;  // events in the open interval
    uint32_t acc, rej, seg;  // per-particle ODE counters (this launch)
};

__device__ __forceinline__ void load_particle(const DevState& D, uint32_t i, Lane& L) {
    const uint32_t o = i * 8u;
    L.pid = i;
    L.K.t = D.ld(F_T, o);
    L.K.t_old = D.ld(F_TOLD, o);
    L.K.next_gas = D.ld(F_NEXTGAS, o);
    L.K.dt = 0.0;
    L.h = D.ld(F_H, o);
    L.tf = D.ld(F_TF, o);
    L.K.pos = d3{D.ld(F_PX, o), D.ld(F_PY, o), D.ld(F_PZ, o)};
    L.K.pos_old = d3{D.ld(F_POX, o), D.ld(F_POY, o), D.ld(F_POZ, o)};
    L.K.v = d3{D.ld(F_VX, o), D.ld(F_VY, o), D.ld(F_VZ, o)};
    L.K.v_old = d3{D.ld(F_VOX, o), D.ld(F_VOY, o), D.ld(F_VOZ, o)};
    L.S = d3{D.ld(F_SX, o), D.ld(F_SY, o), D.ld(F_SZ, o)};
    L.rng.x = D.ld<uint64_t>(F_RX, o);
    L.rng.y = D.ld<uint64_t>(F_RY, o);
    L.rng.z = D.ld<uint64_t>(F_RZ, o);
    L.rng.w = D.ld<uint64_t>(F_RW, o);
    L.rng.spare = D.ld(F_SPARE, o);
    const uint64_t fl = D.ld<uint64_t>(F_FLAGS, o);
    L.K.failure = (int32_t)(uint32_t)fl;
    L.K.stop = (fl >> 32) & 1;
    L.rng.has = (fl >> 33) & 1;
    L.K.coll_type = (char)(uint8_t)(fl >> 40);
    L.K.wall_hit = (char)(uint8_t)(fl >> 48);
    L.K.n_bounce = D.ld<uint64_t>(F_NBOUNCE, o);
    L.K.n_coll = D.ld<uint64_t>(F_NCOLL, o);
    L.n_steps = D.ld<uint64_t>(F_NSTEPS, o);
    L.K.finished = false;
    L.acc = L.rej = L.seg = 0;
}

__device__ __forceinline__ void store_particle(const DevState& D, const Lane& L) {
    const uint32_t o = L.pid * 8u;
    D.st(F_T, o, L.K.t);
    D.st(F_TOLD, o, L.K.t_old);
    D.st(F_NEXTGAS, o, L.K.next_gas);
    D.st(F_H, o, L.h);
    D.st(F_TF, o, L.tf);
    D.st(F_PX, o, L.K.pos.x); D.st(F_PY, o, L.K.pos.y); D.st(F_PZ, o, L.K.pos.z);
    D.st(F_POX, o, L.K.pos_old.x); D.st(F_POY, o, L.K.pos_old.y); D.st(F_POZ, o, L.K.pos_old.z);
    D.st(F_VX, o, L.K.v.x); D.st(F_VY, o, L.K.v.y); D.st(F_VZ, o, L.K.v.z);
    D.st(F_VOX, o, L.K.v_old.x); D.st(F_VOY, o, L.K.v_old.y); D.st(F_VOZ, o, L.K.v_old.z);
    D.st(F_SX, o, L.S.x); D.st(F_SY, o, L.S.y); D.st(F_SZ, o, L.S.z);
    D.st<uint64_t>(F_RX, o, L.rng.x);
    D.st<uint64_t>(F_RY, o, L.rng.y);
    D.st<uint64_t>(F_RZ, o, L.rng.z);
    D.st<uint64_t>(F_RW, o, L.rng.w);
    D.st(F_SPARE, o, L.rng.spare);
    D.st<uint64_t>(F_FLAGS, o, pack_flags(L.K.failure, L.K.stop, L.rng.has, L.K.coll_type, L.K.wall_hit));
    D.st<uint64_t>(F_NBOUNCE, o, L.K.n_bounce);
    D.st<uint64_t>(F_NCOLL, o, L.K.n_coll);
    D.st<uint64_t>(F_NSTEPS, o, L.n_steps);
    D.st<uint64_t>(F_ACC, o, D.ld<uint64_t>(F_ACC, o) + L.acc);
    D.st<uint64_t>(F_REJ, o, D.ld<uint64_t>(F_REJ, o) + L.rej);
    D.st<uint64_t>(F_SEG, o, D.ld<uint64_t>(F_SEG, o) + L.seg);
}

// Snapshot block layout = BinaryLogger::writeSnapshot (src/logger.cpp:14-25):
// t[N] f64 | pos[N] 3xf64 | v[N] | S[N] | failure[N] i32 | n_coll[N] u64 |
// n_bounce[N] u64 | n_steps[N] u64.  The u64 arrays start at 84N, which is
// 8-byte aligned only for even N, so they are written as two u32 halves.
__device__ __forceinline__ void put_u64(char* p, uint64_t v) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    q[0] = (uint32_t)v;
    q[1] = (uint32_t)(v >> 32);
}

__device__ __forceinline__ void write_record(char* B, uint64_t N, uint32_t i, double t, const d3& pos,
                                             const d3& v, const d3& S, int32_t failure, uint64_t n_coll,
                                             uint64_t n_bounce, uint64_t n_steps) {
    reinterpret_cast<double*>(B)[i] = t;
    double* pp = reinterpret_cast<double*>(B + 8 * N) + 3 * (uint64_t)i;
    pp[0] = pos.x; pp[1] = pos.y; pp[2] = pos.z;
    double* pv = reinterpret_cast<double*>(B + 32 * N) + 3 * (uint64_t)i;
    pv[0] = v.x; pv[1] = v.y; pv[2] = v.z;
    double* ps = reinterpret_cast<double*>(B + 56 * N) + 3 * (uint64_t)i;
    ps[0] = S.x; ps[1] = S.y; ps[2] = S.z;
    reinterpret_cast<int32_t*>(B + 80 * N)[i] = failure;
    put_u64(B + 84 * N + 8 * (uint64_t)i, n_coll);
    put_u64(B + 92 * N + 8 * (uint64_t)i, n_bounce);
    put_u64(B + 100 * N + 8 * (uint64_t)i, n_steps);
}

__device__ __forceinline__ double target_of(const RunArgs& R, int32_t k) {
    return R.targets ? R.targets[k] : R.t_single;
}

// Opens the next output interval of the lane's particle, writing the
// snapshot of every interval in which the particle does not move (stopped or
// past its end time).  Returns false when no interval is left.
// src/particle.cpp:748-752: tf = nextTOut; finished = false; tempH = h.
__device__ inline bool open_interval(const RunArgs& R, Lane& L) {
    while (L.k < R.n_targets) {
        const double T = target_of(R, L.k);
        const double prev = (L.k == 0) ? R.t_prev : target_of(R, L.k - 1);
        bool moves = true;
        double tgt = T;
        if (R.stop_at) {
            const double e = R.stop_at[L.pid];
            if (!(prev < e)) moves = false;
            else if (e < T) tgt = e;
        }
        if (moves) {
            L.tf = tgt;
            L.K.finished = false;
            L.tempH = L.h;
            L.n_events = 0;
            if (!L.K.stop) return true;  // the reference's loop runs
        }
        if (R.blocks)
            write_record(R.blocks + (uint64_t)L.k * R.block_bytes, R.n, L.pid, L.K.t, L.K.pos, L.K.v, L.S,
                         L.K.failure, L.K.n_coll, L.K.n_bounce, L.n_steps);
        L.k++;
    }
    return false;
}

__device__ __forceinline__ void close_interval(const RunArgs& R, Lane& L) {
    if (R.blocks)
        write_record(R.blocks + (uint64_t)L.k * R.block_bytes, R.n, L.pid, L.K.t, L.K.pos, L.K.v, L.S,
                     L.K.failure, L.K.n_coll, L.K.n_bounce, L.n_steps);
    L.k++;
}

// Work queue: the lanes that need a particle take consecutive indices with
// one atomic per wave.
__device__ __forceinline__ bool acquire(const RunArgs& R, uint32_t& pid) {
    const uint64_t mask = __ballot(1);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t lane = __lane_id();
    const uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(R.queue, cnt);
    base = __shfl(base, leader);
    const uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    pid = base + below;
    return pid < R.n;
}

// segment end bookkeeping (src/particle.cpp:593-613)
__device__ __forceinline__ void end_segment(const SimParams& P, Lane& L, int status) {
    if (P.keep) L.h = L.tempH;
    if (status < 0) {
        L.K.stop = true;
        L.K.failure = status;
    }
}

__device__ __forceinline__ void poison_and_count(Lane& L) {
    if (L.K.stop) {
        const double nan = __builtin_nan("");
        L.K.t = nan;
        L.K.pos = d3{nan, nan, nan};
        L.K.v = d3{nan, nan, nan};
        L.S = d3{nan, nan, nan};
    }
    L.n_steps += 1;
}

enum { PH_DONE = 0, PH_OPEN = 1, PH_EVENT = 2, PH_INT = 3, PH_CLOSE = 4 };

template <int VAR, int INTEG, bool PERSIST>
__global__ __launch_bounds__(256) void k_engine(const SimParams P, const DevState D, const RunArgs R) {
    Lane L;
    DopState dop;
    RkState rk;
    Seg seg;
    Poly poly;
    uint32_t trips = 0, tacc = 0, trej = 0, tseg = 0;
    int phase = PH_DONE;
    if (PERSIST) {
        uint32_t pid;
        if (acquire(R, pid)) {
            load_particle(D, pid, L);
            L.k = 0;
            phase = PH_OPEN;
        }
    } else {
        const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
        if (pid < R.n) {
            load_particle(D, pid, L);
            L.k = 0;
            phase = PH_OPEN;
        }
    }
    while (phase != PH_DONE) {
        ++trips;
        if (phase == PH_CLOSE) {
            close_interval(R, L);
            phase = PH_OPEN;
        }
        if (phase == PH_OPEN) {
            if (open_interval(R, L)) {
                phase = PH_EVENT;
            } else {
                tacc += L.acc; trej += L.rej; tseg += L.seg;
                store_particle(D, L);
                phase = PH_DONE;
                if (PERSIST) {
                    uint32_t pid;
                    if (acquire(R, pid)) {
                        load_particle(D, pid, L);
                        L.k = 0;
                        phase = PH_OPEN;
                    }
                }
                continue;
            }
        }
        if (phase == PH_EVENT) {
            // one pass of the reference's event loop body up to the
            // integrator call (src/particle.cpp:755-758)
            next_collision(P, L.tf, L.K, L.rng);
            move_and_collide(P, L.K, L.rng);
            sanity(P, L.K);
            if (++L.n_events > R.max_events && L.K.failure == 0) {
                L.K.stop = true;
                L.K.failure = FAIL_WATCHDOG;
            }
            if (L.K.failure == 0 && INTEG != INT_NONE) {
                seg.t0 = L.K.t_old;
                seg.p0 = L.K.pos_old;
                seg.v0 = L.K.v_old;
                if (VAR == VAR_FAST) poly = make_poly(P, seg);
                if (INTEG == INT_DOP853) {
                    if (VAR == VAR_EXACT) dop_begin_exact(P, dop, L.S, seg, L.K.t);
                    else dop_begin_fast(P, dop, L.S, seg, poly, L.K.t);
                } else {
                    rk_begin(rk, seg, L.K.t, L.tempH);
                }
                L.seg++;
                if (R.trace && L.pid == R.trace_pid)
                    trace_record(R, 1, L.K.t_old, L.K.t, L.K.pos_old, L.K.v_old, L.S, L.tempH, L.acc, L.rej,
                                 L.K.coll_type, L.K.wall_hit, 0, L.K.n_bounce, L.K.n_coll);
                phase = PH_INT;
            } else {
                poison_and_count(L);
                phase = (L.K.finished || L.K.stop) ? PH_CLOSE : PH_EVENT;
            }
        }
        if (phase == PH_INT) {
            int st;
            if (INTEG == INT_DOP853) {
                if (VAR == VAR_EXACT) st = dop_trial_exact(P, dop, L.S, L.tempH, seg, L.acc, L.rej);
                else st = dop_trial_fast(P, dop, L.S, L.tempH, seg, poly, L.acc, L.rej);
            } else if (INTEG == INT_RK45) {
                st = rk_trial<VAR>(P, rk, L.S, L.tempH, seg, poly, L.acc, L.rej);
            } else {
                st = 1;  // unreachable: without an integrator no lane enters PH_INT
            }
            if (st != 0) {
                if (R.trace && L.pid == R.trace_pid)
                    trace_record(R, 2, L.K.t_old, L.K.t, L.K.pos_old, L.K.v_old, L.S, L.tempH, L.acc, L.rej,
                                 L.K.coll_type, L.K.wall_hit, st, L.K.n_bounce, L.K.n_coll);
                end_segment(P, L, st > 0 ? 0 : st);
                poison_and_count(L);
                phase = (L.K.finished || L.K.stop) ? PH_CLOSE : PH_EVENT;
            }
        }
    }
    // wave-level reduction of the work counters, then one atomic per wave
    unsigned long long c[4] = {tacc, trej, tseg, trips};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        unsigned long long v = c[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        c[j] = v;
    }
    if (__lane_id() == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) atomicAdd(R.work + j, c[j]);
    }
}

// initParticlesCPU (src/particle.cpp:638-710), one lane per particle; local
// slot j is global particle first + j*stride.
__global__ __launch_bounds__(256) void k_init(const SimParams P, const DevState D, uint32_t n,
                                              uint32_t seed, uint32_t first, uint32_t stride,
                                              double V, double t0, d3 yi, double tf0, int32_t dist) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t ipart = first + j * stride;
    Rng rng;
    seed_rng(rng, seed, ipart);
    d3 p;
    p.x = uni(rng) * P.Lx - P.Lx / 2.0;
    p.y = uni(rng) * P.Ly - P.Ly / 2.0;
    p.z = uni(rng) * P.Lz - P.Lz / 2.0;
    double ng;
    if (P.gas_coll)
        ng = expo_draw(rng, P.tc);
    else
        ng = DBL_MAX_D;
    d3 v{0.0, 0.0, 0.0};
    if (dist == 'C') {
        d3 vec;
        vec.x = normal(rng, 0.0, 1.0);
        vec.y = normal(rng, 0.0, 1.0);
        vec.z = normal(rng, 0.0, 1.0);
        const double vn = len3(vec);
        v = d3{(V * vec.x) / vn, (V * vec.y) / vn, (V * vec.z) / vn};
    } else if (dist == 'M') {
        v.x = maxboltz(rng, P.sqrtKT_m);
        v.y = maxboltz(rng, P.sqrtKT_m);
        v.z = maxboltz(rng, P.sqrtKT_m);
    }
    D.f(F_PX)[j] = p.x; D.f(F_PY)[j] = p.y; D.f(F_PZ)[j] = p.z;
    D.f(F_POX)[j] = p.x; D.f(F_POY)[j] = p.y; D.f(F_POZ)[j] = p.z;
    D.f(F_NEXTGAS)[j] = ng;
    D.f(F_VX)[j] = v.x; D.f(F_VY)[j] = v.y; D.f(F_VZ)[j] = v.z;
    D.f(F_VOX)[j] = v.x; D.f(F_VOY)[j] = v.y; D.f(F_VOZ)[j] = v.z;
    D.f<uint64_t>(F_RX)[j] = rng.x;
    D.f<uint64_t>(F_RY)[j] = rng.y;
    D.f<uint64_t>(F_RZ)[j] = rng.z;
    D.f<uint64_t>(F_RW)[j] = rng.w;
    D.f(F_SPARE)[j] = rng.spare;
    D.f(F_SX)[j] = yi.x; D.f(F_SY)[j] = yi.y; D.f(F_SZ)[j] = yi.z;
    D.f(F_T)[j] = t0;
    D.f(F_TOLD)[j] = t0;
    D.f(F_TF)[j] = tf0;
    D.f(F_H)[j] = P.h0;
    // coll_type 'W', wall_hit 0 (uninitialised in the reference), no failure
    D.f<uint64_t>(F_FLAGS)[j] = pack_flags(0, false, rng.has, 'W', 0);
    D.f<uint64_t>(F_NBOUNCE)[j] = 0;
    D.f<uint64_t>(F_NCOLL)[j] = 0;
    D.f<uint64_t>(F_NSTEPS)[j] = 0;
    D.f<uint64_t>(F_PARTID)[j] = ipart;
    D.f<uint64_t>(F_ACC)[j] = 0;
    D.f<uint64_t>(F_REJ)[j] = 0;
    D.f<uint64_t>(F_SEG)[j] = 0;
}

// Current state -> one snapshot block (reference file layout).
__global__ __launch_bounds__(256) void k_snapshot(const DevState D, uint32_t n, char* B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    write_record(B, n, i, D.f(F_T)[i], d3{D.f(F_PX)[i], D.f(F_PY)[i], D.f(F_PZ)[i]},
                 d3{D.f(F_VX)[i], D.f(F_VY)[i], D.f(F_VZ)[i]}, d3{D.f(F_SX)[i], D.f(F_SY)[i], D.f(F_SZ)[i]},
                 (int32_t)(uint32_t)D.f<uint64_t>(F_FLAGS)[i], D.f<uint64_t>(F_NCOLL)[i],
                 D.f<uint64_t>(F_NBOUNCE)[i], D.f<uint64_t>(F_NSTEPS)[i]);
}

}  // namespace st

// ===========================================================================
// C ABI
// ===========================================================================
using namespace st;

namespace {

thread_local std::string g_create_error;

struct Buffers {
    void* base = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct st_handle {
    st_options opt;
    st_ext ext;
    SimParams P;
    DevState D;
    uint32_t n = 0;
    Buffers state;
    double* d_targets = nullptr;
    int32_t targets_cap = 0;
    double* d_stop_at = nullptr;
    bool have_stop_at = false;
    unsigned long long* d_work = nullptr;
    uint32_t* d_queue = nullptr;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool initialized = false;
    double last_target = 0.0;
    int persist_grid = 0;
    // segment trace of one particle
    double* d_trace = nullptr;
    uint32_t* d_trace_count = nullptr;
    uint32_t trace_cap = 0;
    int64_t trace_pid = -1;
    std::string err;
};

static int fail(st_handle* h, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIPCHK(h, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((h), ST_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));        \
    } while (0)

static SimParams derive_params(const st_options& o, const st_ext& x) {
    SimParams P;
    memset(&P, 0, sizeof P);
    P.Lx = o.L.x; P.Ly = o.L.y; P.Lz = o.L.z;
    P.B0x = o.B0.x; P.B0y = o.B0.y; P.B0z = o.B0.z;
    P.Ex = o.E.x; P.Ey = o.E.y; P.Ez = o.E.z;
    P.G[0] = o.Gx.x; P.G[1] = o.Gx.y; P.G[2] = o.Gx.z;
    P.G[3] = o.Gy.x; P.G[4] = o.Gy.y; P.G[5] = o.Gy.z;
    P.G[6] = o.Gz.x; P.G[7] = o.Gz.y; P.G[8] = o.Gz.z;
    P.gamma = o.gamma;
    P.pa = o.a;
    P.pw = o.w;
    P.pulse_axis = x.pulse_axis;
    const double b0ax = x.pulse_axis == 0 ? o.B0.x : (x.pulse_axis == 1 ? o.B0.y : o.B0.z);
    // a*cos(w*t) with a == 0 is a signed zero; added to B0[axis] it changes
    // nothing unless B0[axis] is itself -0.0 (or cos is not finite).
    P.pulse_on = (o.a != 0.0 || signbit(b0ax) || !isfinite(o.w)) ? 1 : 0;
    P.rtol = o.rtol; P.atol = o.atol; P.beta = o.beta; P.safe = o.safe;
    P.fac1 = o.fac1; P.fac2 = o.fac2; P.hmax = o.hmax; P.hmin = o.hmin; P.h0 = o.h;
    P.expo1 = 1.0 / 8.0 - o.beta * 0.2;
    P.facc1 = 1.0 / o.fac1;
    P.facc2 = 1.0 / o.fac2;
    P.pow_beta_on = (o.beta != 0.0) ? 1 : 0;
    P.maxPosStep = o.maxPosStep;
    P.tc = o.tc;
    P.sqrtKT_m = o.sqrtKT_m;
    P.diffuse = o.diffuse;
    P.diffuse_on = (o.diffuse > (double)1.17549435082228750797e-38) ? 1 : 0;
    P.nmax = o.nmax;
    P.dist = (int32_t)(unsigned char)o.dist;
    P.gravity = o.gravity ? 1 : 0;
    P.gas_coll = o.gas_coll ? 1 : 0;
    P.fixed = o.fixedStepSize ? 1 : 0;
    P.keep = o.keepStepSize ? 1 : 0;
    return P;
}

static uint64_t field_stride(uint32_t n) { return ((uint64_t)8 * n + 255) & ~(uint64_t)255; }

template <int VAR, int INTEG, bool PERSIST>
static hipError_t launch_engine(st_handle* h, dim3 grid, dim3 block, const RunArgs& R) {
    hipLaunchKernelGGL((k_engine<VAR, INTEG, PERSIST>), grid, block, 0, h->stream, h->P, h->D, R);
    return hipGetLastError();
}

using EngineFn = hipError_t (*)(st_handle*, dim3, dim3, const RunArgs&);

template <bool PERSIST>
static EngineFn pick_engine(int variant, int integ) {
    if (variant == ST_VARIANT_FAST) {
        if (integ == 0) return launch_engine<VAR_FAST, INT_DOP853, PERSIST>;
        if (integ == 1) return launch_engine<VAR_FAST, INT_RK45, PERSIST>;
        return launch_engine<VAR_FAST, INT_NONE, PERSIST>;
    }
    if (integ == 0) return launch_engine<VAR_EXACT, INT_DOP853, PERSIST>;
    if (integ == 1) return launch_engine<VAR_EXACT, INT_RK45, PERSIST>;
    return launch_engine<VAR_EXACT, INT_NONE, PERSIST>;
}

template <int VAR, int INTEG>
static const void* engine_symbol() {
    return reinterpret_cast<const void*>(&k_engine<VAR, INTEG, true>);
}

static const void* persist_symbol(int variant, int integ) {
    if (variant == ST_VARIANT_FAST) {
        if (integ == 0) return engine_symbol<VAR_FAST, INT_DOP853>();
        if (integ == 1) return engine_symbol<VAR_FAST, INT_RK45>();
        return engine_symbol<VAR_FAST, INT_NONE>();
    }
    if (integ == 0) return engine_symbol<VAR_EXACT, INT_DOP853>();
    if (integ == 1) return engine_symbol<VAR_EXACT, INT_RK45>();
    return engine_symbol<VAR_EXACT, INT_NONE>();
}

static RunArgs base_args(st_handle* h) {
    RunArgs R;
    memset(&R, 0, sizeof R);
    R.n = h->n;
    R.work = h->d_work;
    R.queue = h->d_queue;
    R.stop_at = h->have_stop_at ? h->d_stop_at : nullptr;
    R.block_bytes = (uint64_t)108 * h->n;
    R.max_events = 100000000u;
    if (h->trace_pid >= 0 && h->d_trace) {
        R.trace = h->d_trace;
        R.trace_count = h->d_trace_count;
        R.trace_pid = (uint32_t)h->trace_pid;
        R.trace_cap = h->trace_cap;
    }
    return R;
}

static int block_size(const st_handle* h) {
    int b = h->ext.block_size > 0 ? h->ext.block_size : 256;
    if (b > 256) b = 256;
    if (b % 64) b = (b / 64 + 1) * 64;
    return b;
}

extern "C" {

ST_API const char* st_version(void) { return "spintrack gfx950 exact,fast dop853,rk45"; }

ST_API const char* st_last_error(const st_handle* h) {
    return h ? h->err.c_str() : g_create_error.c_str();
}

ST_API int st_create(const st_options* opt, const st_ext* ext, st_handle** out) {
    if (!opt || !out) return fail(nullptr, ST_E_ARG, "st_create: null argument");
    *out = nullptr;
    if (opt->numParticles <= 0) return fail(nullptr, ST_E_ARG, "numParticles must be > 0 (got %d)", opt->numParticles);
    if (opt->integratorType > 1 || opt->integratorType < -1)
        return fail(nullptr, ST_E_UNSUPPORTED,
                    "integratorType %d is not implemented (0 = DOP853, 1 = RK45, -1 = trajectory only)",
                    opt->integratorType);
    st_ext x{2, ST_VARIANT_EXACT, 0, 0};
    if (ext) x = *ext;
    if (x.pulse_axis < 0 || x.pulse_axis > 2) return fail(nullptr, ST_E_ARG, "pulse_axis must be 0, 1 or 2");
    if (x.variant != ST_VARIANT_EXACT && x.variant != ST_VARIANT_FAST)
        return fail(nullptr, ST_E_ARG, "unknown variant %d", x.variant);
    st_handle* h = new (std::nothrow) st_handle();
    if (!h) return fail(nullptr, ST_E_ARG, "out of host memory");
    h->opt = *opt;
    h->ext = x;
    h->n = (uint32_t)opt->numParticles;
    h->P = derive_params(*opt, x);
    const size_t bytes = (size_t)field_stride(h->n) * F_COUNT;
    hipError_t e = hipMalloc(&h->state.base, bytes);
    if (e != hipSuccess) {
        fail(nullptr, ST_E_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        delete h;
        return ST_E_HIP;
    }
    h->state.bytes = bytes;
    h->D.base = (char*)h->state.base;
    h->D.stride = field_stride(h->n);
    if ((e = hipMalloc(&h->d_work, 4 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&h->d_queue, 256)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking)) != hipSuccess) {
        fail(nullptr, ST_E_HIP, "device set-up failed: %s", hipGetErrorString(e));
        st_destroy(h);
        return ST_E_HIP;
    }
    h->stream = h->own_stream;
    hipMemsetAsync(h->d_work, 0, 4 * sizeof(unsigned long long), h->stream);
    *out = h;
    return ST_OK;
}

ST_API void st_destroy(st_handle* h) {
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->state.base) hipFree(h->state.base);
    if (h->d_targets) hipFree(h->d_targets);
    if (h->d_stop_at) hipFree(h->d_stop_at);
    if (h->d_work) hipFree(h->d_work);
    if (h->d_queue) hipFree(h->d_queue);
    if (h->d_trace) hipFree(h->d_trace);
    if (h->d_trace_count) hipFree(h->d_trace_count);
    if (h->own_stream) hipStreamDestroy(h->own_stream);
    delete h;
}

ST_API int st_set_stream(st_handle* h, void* s) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    h->stream = s ? (hipStream_t)s : h->own_stream;
    return ST_OK;
}

ST_API int st_init_subset(st_handle* h, uint32_t first, uint32_t stride) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    const st_options& o = h->opt;
    const int bs = 256;
    const dim3 grid((h->n + bs - 1) / bs);
    hipLaunchKernelGGL(k_init, grid, dim3(bs), 0, h->stream, h->P, h->D, h->n, o.seed, first, stride,
                       o.V, o.t0, d3{o.yi.x, o.yi.y, o.yi.z}, o.tf, (int32_t)(unsigned char)o.dist);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemsetAsync(h->d_work, 0, 4 * sizeof(unsigned long long), h->stream));
    h->initialized = true;
    h->last_target = o.t0;
    return ST_OK;
}

ST_API int st_init(st_handle* h) { return st_init_subset(h, 0, 1); }

ST_API int st_set_stop_times(st_handle* h, const double* stop_at) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (!stop_at) {
        h->have_stop_at = false;
        return ST_OK;
    }
    if (!h->d_stop_at) HIPCHK(h, hipMalloc(&h->d_stop_at, sizeof(double) * h->n));
    HIPCHK(h, hipMemcpyAsync(h->d_stop_at, stop_at, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_stop_at = true;
    return ST_OK;
}

ST_API int st_advance_to(st_handle* h, double next_t) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (!h->initialized) return fail(h, ST_E_STATE, "st_advance_to before st_init");
    RunArgs R = base_args(h);
    R.targets = nullptr;
    R.t_single = next_t;
    R.t_prev = h->last_target;
    R.n_targets = 1;
    const int bs = block_size(h);
    const dim3 grid((h->n + bs - 1) / bs);
    hipError_t e = pick_engine<false>(h->ext.variant, h->opt.integratorType)(h, grid, dim3(bs), R);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "engine launch failed: %s", hipGetErrorString(e));
    h->last_target = next_t;
    return ST_OK;
}

ST_API int st_run(st_handle* h, int32_t n_targets, const double* targets, void* blocks_dev) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (!h->initialized) return fail(h, ST_E_STATE, "st_run before st_init");
    if (n_targets <= 0 || !targets) return fail(h, ST_E_ARG, "st_run needs at least one target");
    for (int k = 1; k < n_targets; ++k)
        if (!(targets[k] > targets[k - 1])) return fail(h, ST_E_ARG, "targets must increase (index %d)", k);
    if (n_targets > h->targets_cap) {
        if (h->d_targets) HIPCHK(h, hipFree(h->d_targets));
        h->d_targets = nullptr;
        HIPCHK(h, hipMalloc(&h->d_targets, sizeof(double) * n_targets));
        h->targets_cap = n_targets;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_targets, targets, sizeof(double) * n_targets, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_queue, 0, sizeof(uint32_t), h->stream));
    if (h->persist_grid == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        HIPCHK(h, hipGetDevice(&dev));
        HIPCHK(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &per_cu, persist_symbol(h->ext.variant, h->opt.integratorType), 256, 0));
        if (per_cu < 1) per_cu = 1;
        h->persist_grid = cus * per_cu;
    }
    RunArgs R = base_args(h);
    R.targets = h->d_targets;
    R.t_prev = h->last_target;
    R.n_targets = n_targets;
    R.blocks = (char*)blocks_dev;
    int grid = h->persist_grid;
    const int need = (int)((h->n + 255) / 256);
    if (grid > need) grid = need;
    hipError_t e = pick_engine<true>(h->ext.variant, h->opt.integratorType)(h, dim3(grid), dim3(256), R);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "engine launch failed: %s", hipGetErrorString(e));
    h->last_target = targets[n_targets - 1];
    return ST_OK;
}

ST_API int st_synchronize(st_handle* h) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ST_OK;
}

ST_API size_t st_snapshot_block_bytes(const st_handle* h) { return h ? (size_t)108 * h->n : 0; }

ST_API int st_snapshot_block_device(st_handle* h, void* dst) {
    if (!h || !dst) return fail(h, ST_E_ARG, "null argument");
    const int bs = 256;
    hipLaunchKernelGGL(k_snapshot, dim3((h->n + bs - 1) / bs), dim3(bs), 0, h->stream, h->D, h->n, (char*)dst);
    HIPCHK(h, hipGetLastError());
    return ST_OK;
}

// Copies the whole SoA state to host memory (one transfer).
static int state_to_host(st_handle* h, std::vector<char>& host) {
    host.resize(h->state.bytes);
    HIPCHK(h, hipMemcpyAsync(host.data(), h->state.base, h->state.bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ST_OK;
}

ST_API int st_snapshot(st_handle* h, double* t, st_coords* pos, st_coords* v, st_coords* S, int32_t* failure,
                       uint64_t* n_coll, uint64_t* n_bounce, uint64_t* n_steps) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    std::vector<char> host;
    int r = state_to_host(h, host);
    if (r) return r;
    const DevState H{host.data(), h->D.stride};
    for (uint32_t i = 0; i < h->n; ++i) {
        if (t) t[i] = H.f(F_T)[i];
        if (pos) pos[i] = st_coords{H.f(F_PX)[i], H.f(F_PY)[i], H.f(F_PZ)[i]};
        if (v) v[i] = st_coords{H.f(F_VX)[i], H.f(F_VY)[i], H.f(F_VZ)[i]};
        if (S) S[i] = st_coords{H.f(F_SX)[i], H.f(F_SY)[i], H.f(F_SZ)[i]};
        if (failure) failure[i] = (int32_t)(uint32_t)H.f<uint64_t>(F_FLAGS)[i];
        if (n_coll) n_coll[i] = H.f<uint64_t>(F_NCOLL)[i];
        if (n_bounce) n_bounce[i] = H.f<uint64_t>(F_NBOUNCE)[i];
        if (n_steps) n_steps[i] = H.f<uint64_t>(F_NSTEPS)[i];
    }
    return ST_OK;
}

ST_API int st_get_state(st_handle* h, st_particle* out) {
    if (!h || !out) return fail(h, ST_E_ARG, "null argument");
    std::vector<char> host;
    int r = state_to_host(h, host);
    if (r) return r;
    const DevState H{host.data(), h->D.stride};
    for (uint32_t i = 0; i < h->n; ++i) {
        st_particle& q = out[i];
        memset(&q, 0, sizeof q);
        q.S = st_coords{H.f(F_SX)[i], H.f(F_SY)[i], H.f(F_SZ)[i]};
        q.v = st_coords{H.f(F_VX)[i], H.f(F_VY)[i], H.f(F_VZ)[i]};
        q.v_old = st_coords{H.f(F_VOX)[i], H.f(F_VOY)[i], H.f(F_VOZ)[i]};
        q.pos = st_coords{H.f(F_PX)[i], H.f(F_PY)[i], H.f(F_PZ)[i]};
        q.pos_old = st_coords{H.f(F_POX)[i], H.f(F_POY)[i], H.f(F_POZ)[i]};
        q.t = H.f(F_T)[i]; q.t_old = H.f(F_TOLD)[i]; q.tf = H.f(F_TF)[i];
        q.next_gas = H.f(F_NEXTGAS)[i]; q.h = H.f(F_H)[i];
        q.rng.x = H.f<uint64_t>(F_RX)[i]; q.rng.y = H.f<uint64_t>(F_RY)[i];
        q.rng.z = H.f<uint64_t>(F_RZ)[i]; q.rng.w = H.f<uint64_t>(F_RW)[i];
        q.rng.spare = H.f(F_SPARE)[i];
        const uint64_t fl = H.f<uint64_t>(F_FLAGS)[i];
        q.failure = (int32_t)(uint32_t)fl;
        q.stop = (fl >> 32) & 1;
        q.rng.hasSpare = (fl >> 33) & 1;
        q.coll_type = (uint8_t)(fl >> 40);
        q.wall_hit = (uint8_t)(fl >> 48);
        q.n_bounce = H.f<uint64_t>(F_NBOUNCE)[i]; q.n_coll = H.f<uint64_t>(F_NCOLL)[i];
        q.n_steps = H.f<uint64_t>(F_NSTEPS)[i];
        q.part_id = (uint32_t)H.f<uint64_t>(F_PARTID)[i];
    }
    return ST_OK;
}

ST_API int st_set_state(st_handle* h, const st_particle* in) {
    if (!h || !in) return fail(h, ST_E_ARG, "null argument");
    std::vector<char> host;
    int r = state_to_host(h, host);
    if (r) return r;
    const DevState H{host.data(), h->D.stride};
    for (uint32_t i = 0; i < h->n; ++i) {
        const st_particle& q = in[i];
        H.f(F_SX)[i] = q.S.x; H.f(F_SY)[i] = q.S.y; H.f(F_SZ)[i] = q.S.z;
        H.f(F_VX)[i] = q.v.x; H.f(F_VY)[i] = q.v.y; H.f(F_VZ)[i] = q.v.z;
        H.f(F_VOX)[i] = q.v_old.x; H.f(F_VOY)[i] = q.v_old.y; H.f(F_VOZ)[i] = q.v_old.z;
        H.f(F_PX)[i] = q.pos.x; H.f(F_PY)[i] = q.pos.y; H.f(F_PZ)[i] = q.pos.z;
        H.f(F_POX)[i] = q.pos_old.x; H.f(F_POY)[i] = q.pos_old.y; H.f(F_POZ)[i] = q.pos_old.z;
        H.f(F_T)[i] = q.t; H.f(F_TOLD)[i] = q.t_old; H.f(F_TF)[i] = q.tf;
        H.f(F_NEXTGAS)[i] = q.next_gas; H.f(F_H)[i] = q.h;
        H.f<uint64_t>(F_RX)[i] = q.rng.x; H.f<uint64_t>(F_RY)[i] = q.rng.y;
        H.f<uint64_t>(F_RZ)[i] = q.rng.z; H.f<uint64_t>(F_RW)[i] = q.rng.w;
        H.f(F_SPARE)[i] = q.rng.spare;
        H.f<uint64_t>(F_FLAGS)[i] = pack_flags(q.failure, q.stop != 0, q.rng.hasSpare != 0, (char)q.coll_type,
                                               (char)q.wall_hit);
        H.f<uint64_t>(F_NBOUNCE)[i] = q.n_bounce; H.f<uint64_t>(F_NCOLL)[i] = q.n_coll;
        H.f<uint64_t>(F_NSTEPS)[i] = q.n_steps;
        H.f<uint64_t>(F_PARTID)[i] = q.part_id;
    }
    HIPCHK(h, hipMemcpyAsync(h->state.base, host.data(), h->state.bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->initialized = true;
    return ST_OK;
}

ST_API int st_particle_counters(st_handle* h, uint64_t* accepted, uint64_t* rejected, uint64_t* segments) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    std::vector<char> host;
    int r = state_to_host(h, host);
    if (r) return r;
    const DevState H{host.data(), h->D.stride};
    for (uint32_t i = 0; i < h->n; ++i) {
        if (accepted) accepted[i] = H.f<uint64_t>(F_ACC)[i];
        if (rejected) rejected[i] = H.f<uint64_t>(F_REJ)[i];
        if (segments) segments[i] = H.f<uint64_t>(F_SEG)[i];
    }
    return ST_OK;
}

ST_API int st_set_trace(st_handle* h, int64_t particle, uint32_t capacity) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (particle < 0) {
        h->trace_pid = -1;
        return ST_OK;
    }
    if ((uint64_t)particle >= h->n) return fail(h, ST_E_ARG, "trace particle %lld out of range", (long long)particle);
    if (capacity == 0) return fail(h, ST_E_ARG, "trace capacity must be > 0");
    if (capacity > h->trace_cap) {
        if (h->d_trace) HIPCHK(h, hipFree(h->d_trace));
        h->d_trace = nullptr;
        HIPCHK(h, hipMalloc(&h->d_trace, sizeof(double) * TRACE_DOUBLES * (size_t)capacity));
        h->trace_cap = capacity;
    }
    if (!h->d_trace_count) HIPCHK(h, hipMalloc(&h->d_trace_count, sizeof(uint32_t)));
    HIPCHK(h, hipMemsetAsync(h->d_trace_count, 0, sizeof(uint32_t), h->stream));
    h->trace_pid = particle;
    return ST_OK;
}

ST_API int st_get_trace(st_handle* h, double* out, uint32_t capacity, uint32_t* n_records) {
    if (!h || !n_records) return fail(h, ST_E_ARG, "null argument");
    uint32_t n = 0;
    if (h->d_trace_count) {
        HIPCHK(h, hipMemcpyAsync(&n, h->d_trace_count, sizeof n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (n > h->trace_cap) n = h->trace_cap;
    *n_records = n;
    const uint32_t m = n < capacity ? n : capacity;
    if (out && m) {
        HIPCHK(h, hipMemcpyAsync(out, h->d_trace, sizeof(double) * TRACE_DOUBLES * (size_t)m, hipMemcpyDeviceToHost,
                                 h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return ST_OK;
}

ST_API int st_work_counters(st_handle* h, uint64_t out[4]) {
    if (!h || !out) return fail(h, ST_E_ARG, "null argument");
    unsigned long long w[4];
    HIPCHK(h, hipMemcpyAsync(w, h->d_work, sizeof w, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int j = 0; j < 4; ++j) out[j] = w[j];
    return ST_OK;
}

}  // extern "C"
