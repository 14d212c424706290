// st_state.h -- device-resident particle state, snapshot blocks and the
// per-lane particle record shared by the lane kernels.
#pragma once

#include "st_integrators.h"

namespace st {

// ---------------------------------------------------------------------------
// Particle state: one allocation, one 8-byte array per field; field f of
// particle i lives at base + f*stride + 8*i.  Passing one base pointer
// instead of ~40 array pointers keeps the kernel arguments small.
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
    // pipeline bookkeeping (kinematics kernel / spin kernel cursors)
    F_KSTATE,     // kinematics: open interval index (low 32) | interval-open flag <<32 | done <<33
    F_SSTATE,     // spin: interval index (low 32) | done << 33
    F_TEMPH,      // spin: step size carried between segments of an interval
    F_TSPIN,      // spin: end time of the last integrated segment
    F_NEVENTS,    // kinematics: events in the open interval
    F_CFG,        // batched handles: configuration index of the particle
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
    // alive from the load at the top of a loop to the store at its end.
    template <typename T = double>
    __device__ __forceinline__ T ld(int field, uint32_t off) const {
        return *reinterpret_cast<const T*>(base + (uint64_t)field * stride + off);
    }
    template <typename T = double>
    __device__ __forceinline__ void st(int field, uint32_t off, T v) const {
        *reinterpret_cast<T*>(base + (uint64_t)field * stride + off) = v;
    }
};

ST_HD __forceinline__ uint64_t pack_flags(int32_t failure, bool stop, bool has_spare, char ct, char wh) {
    return (uint64_t)(uint32_t)failure | ((uint64_t)(stop ? 1 : 0) << 32) | ((uint64_t)(has_spare ? 1 : 0) << 33) |
           ((uint64_t)(uint8_t)ct << 40) | ((uint64_t)(uint8_t)wh << 48);
}

// ---------------------------------------------------------------------------
// Snapshot block = BinaryLogger::writeSnapshot layout (src/logger.cpp:14-25):
// t[N] f64 | pos[N] 3xf64 | v[N] | S[N] | failure[N] i32 | n_coll[N] u64 |
// n_bounce[N] u64 | n_steps[N] u64 -- 108*N bytes.  The u64 arrays start at
// 84N, 8-byte aligned only for even N, so they are written as u32 halves.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put_u64(char* p, uint64_t v) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    q[0] = (uint32_t)v;
    q[1] = (uint32_t)(v >> 32);
}

// the fields the kinematics side owns
__device__ __forceinline__ void write_block_kin(char* B, uint64_t N, uint32_t i, double t, const d3& pos,
                                                const d3& v, int32_t failure, uint64_t n_coll, uint64_t n_bounce,
                                                uint64_t n_steps) {
    reinterpret_cast<double*>(B)[i] = t;
    double* pp = reinterpret_cast<double*>(B + 8 * N) + 3 * (uint64_t)i;
    pp[0] = pos.x; pp[1] = pos.y; pp[2] = pos.z;
    double* pv = reinterpret_cast<double*>(B + 32 * N) + 3 * (uint64_t)i;
    pv[0] = v.x; pv[1] = v.y; pv[2] = v.z;
    reinterpret_cast<int32_t*>(B + 80 * N)[i] = failure;
    put_u64(B + 84 * N + 8 * (uint64_t)i, n_coll);
    put_u64(B + 92 * N + 8 * (uint64_t)i, n_bounce);
    put_u64(B + 100 * N + 8 * (uint64_t)i, n_steps);
}

// the field the spin side owns
__device__ __forceinline__ void write_block_spin(char* B, uint64_t N, uint32_t i, const d3& S) {
    double* ps = reinterpret_cast<double*>(B + 56 * N) + 3 * (uint64_t)i;
    ps[0] = S.x; ps[1] = S.y; ps[2] = S.z;
}

__device__ __forceinline__ void write_record(char* B, uint64_t N, uint32_t i, double t, const d3& pos,
                                             const d3& v, const d3& S, int32_t failure, uint64_t n_coll,
                                             uint64_t n_bounce, uint64_t n_steps) {
    write_block_kin(B, N, i, t, pos, v, failure, n_coll, n_bounce, n_steps);
    write_block_spin(B, N, i, S);
}

// ---------------------------------------------------------------------------
// Launch arguments shared by the lane kernels.
// ---------------------------------------------------------------------------
struct RunArgs {
    const double* targets;   // device array of output times, or nullptr
    double t_single;         // the only target when targets == nullptr
    double t_prev;           // the target before targets[0]
    const double* stop_at;   // per-particle end times or nullptr
    char* blocks;            // snapshot blocks or nullptr
    uint64_t block_bytes;
    unsigned long long* work;  // counters: accepted, rejected, segments, trips, events, kin trips
    uint32_t* queue;         // work-queue head (persistent kernels)
    uint32_t n;              // particles
    int32_t n_targets;
    uint32_t max_events;     // per-interval event cap (watchdog against runaway lanes)
    // segment trace of one particle (trace_pid), or trace == nullptr
    double* trace;
    uint32_t* trace_count;
    uint32_t trace_pid;
    uint32_t trace_cap;
};

constexpr int32_t FAIL_WATCHDOG = -9;   // ST_FAIL_WATCHDOG of the C ABI

__device__ __forceinline__ double target_of(const RunArgs& R, int32_t k) {
    return R.targets ? R.targets[k] : R.t_single;
}

// Trace records (TRACE_DOUBLES doubles each), two per integrated segment:
// tag (1 = segment start, 2 = segment end), t_old, t, pos_old[3], v_old[3],
// S[3], step size, accepted, rejected (this particle, this launch),
// coll_type, wall_hit, integrator status, n_bounce, n_coll.
constexpr int TRACE_DOUBLES = 20;

__device__ __forceinline__ void trace_put(double* trace, uint32_t* count, uint32_t cap, int tag, double t_old,
                                       double t, d3 po, d3 vo, d3 s, double h, uint32_t acc, uint32_t rej, char ct,
                                       char wh, int status, uint64_t nb, uint64_t nc) {
    const uint32_t k = atomicAdd(count, 1u);
    if (k >= cap) return;
    double* r = trace + (uint64_t)k * TRACE_DOUBLES;
    r[0] = tag; r[1] = t_old; r[2] = t;
    r[3] = po.x; r[4] = po.y; r[5] = po.z;
    r[6] = vo.x; r[7] = vo.y; r[8] = vo.z;
    r[9] = s.x; r[10] = s.y; r[11] = s.z;
    r[12] = h; r[13] = acc; r[14] = rej; r[15] = (double)ct; r[16] = (double)wh;
    r[17] = status; r[18] = (double)nb; r[19] = (double)nc;
}

#define ST_TRACE(R, ...) trace_put((R).trace, (R).trace_count, (R).trace_cap, __VA_ARGS__)

// Work queue: the lanes that need a particle take consecutive indices with
// one atomic per wave.
__device__ __forceinline__ uint32_t acquire_index(uint32_t* queue) {
    const uint64_t mask = __ballot(1);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t lane = __lane_id();
    const uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(queue, cnt);
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Wave reduction of per-lane counters, one atomic per wave; call with the
// whole wave converged.
template <int N>
__device__ __forceinline__ void flush_counters(unsigned long long* dst, unsigned long long (&c)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        unsigned long long v = c[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        c[j] = v;
    }
    if (__lane_id() == 0) {
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (c[j]) atomicAdd(dst + j, c[j]);
    }
}

}  // namespace st
