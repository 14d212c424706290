// st_fused.h -- the single-kernel lane engine: kinematics and spin
// integration interleaved in one flat loop per lane (one event and one
// integrator trial per trip).  Kept beside the two-kernel pipeline
// (st_pipeline.h) as an alternative engine for comparison; both produce
// identical per-particle results.
#pragma once

#include "st_state.h"

namespace st {

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

// PULSE = false: the fast DOP853 trial's pulse-free form (handles without a
// pulse term), the same form the pipeline's spin kernel takes for them
template <int VAR, int INTEG, bool PERSIST, bool PULSE = true>
__global__ __launch_bounds__(256) void k_engine(const SimParams P, const DevState D, const RunArgs R) {
    const Phys& F = P.ph;
    Lane L;
    DopState dop;
    RkState rk;
    Seg seg;
    Poly poly;
    uint32_t trips = 0, tacc = 0, trej = 0, tseg = 0;
    int phase = PH_DONE;
    if (PERSIST) {
        const uint32_t pid = acquire_index(R.queue);
        if (pid < R.n) {
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
                    const uint32_t pid = acquire_index(R.queue);
                    if (pid < R.n) {
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
            next_collision(P, F, L.tf, L.K, L.rng);
            move_and_collide(P, F, L.K, L.rng);
            sanity(P, L.K);
            if (++L.n_events > R.max_events && L.K.failure == 0) {
                L.K.stop = true;
                L.K.failure = FAIL_WATCHDOG;
            }
            if (L.K.failure == 0 && INTEG != INT_NONE) {
                seg.t0 = L.K.t_old;
                seg.p0 = L.K.pos_old;
                seg.v0 = L.K.v_old;
                if (VAR == VAR_FAST) poly = make_poly(P, F, seg);
                if (INTEG == INT_DOP853) {
                    if (VAR == VAR_EXACT) dop_begin_exact(P, F, dop, L.S, seg, L.K.t);
                    else dop_begin_fast(P, F, dop, L.S, seg, poly, L.K.t);
                } else {
                    rk_begin(rk, seg, L.K.t, L.tempH);
                }
                L.seg++;
                if (R.trace && L.pid == R.trace_pid)
                    ST_TRACE(R, 1, L.K.t_old, L.K.t, L.K.pos_old, L.K.v_old, L.S, L.tempH, L.acc, L.rej,
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
                if (VAR == VAR_EXACT) st = dop_trial_exact(P, F, dop, L.S, L.tempH, seg, L.acc, L.rej);
                else st = dop_trial_fast<PULSE>(P, F, dop, L.S, L.tempH, seg, poly, L.acc, L.rej);
            } else if (INTEG == INT_RK45) {
                st = rk_trial<VAR>(P, F, rk, L.S, L.tempH, seg, poly, L.acc, L.rej);
            } else {
                st = 1;  // unreachable: without an integrator no lane enters PH_INT
            }
            if (st != 0) {
                if (R.trace && L.pid == R.trace_pid)
                    ST_TRACE(R, 2, L.K.t_old, L.K.t, L.K.pos_old, L.K.v_old, L.S, L.tempH, L.acc, L.rej,
                                 L.K.coll_type, L.K.wall_hit, st, L.K.n_bounce, L.K.n_coll);
                end_segment(P, L, st > 0 ? 0 : st);
                poison_and_count(L);
                phase = (L.K.finished || L.K.stop) ? PH_CLOSE : PH_EVENT;
            }
        }
    }
    unsigned long long c[4] = {tacc, trej, tseg, trips};
    flush_counters(R.work, c);
}

}  // namespace st
