// st_pipeline.h -- the two-kernel lane pipeline.
//
// The particle's motion never depends on its spin (the fields do not act
// back on the trajectory), so the reference's event loop
// (src/particle.cpp:754-803) separates into
//   * k_kin:  the branchy kinematics -- collision times, moves, wall/gas
//             reflections, RNG, sanity checks (src/particle.cpp:121-444) --
//             which writes one 64-byte record per trajectory segment to HBM;
//   * k_spin: the uniform FP64 work -- one DOP853/RK45 trial step per loop
//             trip (src/integrator.cpp:90-337) -- which reads the records.
// Every lane of a k_kin wave handles one event per trip, and a k_spin lane
// only pays a record load and a segment set-up between trial steps, so
// neither kernel runs the other's code on idle lanes.  Runs are cut into
// chunks of at most C records per particle (C from the staging-buffer
// budget); particles that finish drop out of the next chunk's active list.
//
// A spin-integrator failure (DOP853 nstep > nmax, RK45 h < hmin or nstep >
// nmax) stops the reference's event loop at that segment, so k_fixup rewinds
// the particle's kinematics to the chunk start and replays it exactly up to
// the failing segment before applying the stop (rare path).
#pragma once

#include "st_state.h"

namespace st {

// record flags
constexpr uint64_t R_INTEGRATE = 1u;     // integrate the spin over this segment
constexpr uint64_t R_IEND = 2u;          // the open interval closes after this record
constexpr uint64_t R_STOP = 4u;          // the particle stops after this record
constexpr uint64_t R_FIRST = 8u;         // first segment of an interval: tempH = h
constexpr uint64_t R_EMPTY = 16u;        // interval closes without any event
// bits 8-15 coll_type, 16-23 wall_hit, 32-63 failure code applied at R_STOP

struct SegRec {
    double t;              // segment end time (t after the move)
    double px, py, pz;     // pos_old: position at the segment start
    double vx, vy, vz;     // v_old: velocity during the segment (pre-move)
    uint64_t flags;
};
static_assert(sizeof(SegRec) == 64, "segment record is one 64-byte line");

struct PipeArgs {
    RunArgs R;
    char* recs;              // N * C records
    uint32_t C;              // record capacity per particle per chunk
    const uint32_t* list;    // active particles this chunk
    uint32_t n_list;
    uint32_t* next_list;     // particles still active after this chunk
    uint32_t* next_count;
    uint32_t* seg_count;     // records written per particle this chunk
    char* backup;            // kinematic state at chunk start (for k_fixup)
    uint32_t* fail_list;     // (pid, record index, status) triples
    uint32_t* fail_count;
    uint32_t fail_cap;
    int32_t integ;           // integratorType
    const Phys* phys;        // batched handles: physics per configuration (F_CFG indexes it)
    // graph-captured chunks: the list length is read from here (set on the
    // device by the chunk head) instead of n_list; nullptr otherwise
    const uint32_t* n_list_dev;
    // Overflow pages: the records j >= C of a particle that is near its end
    // when it reaches the chunk depth, in pages of OVF_PAGE records
    // linked per particle, taken from a shared pool.  The number of segments
    // per particle is heavy-tailed (a few bounce far more often than the
    // mean), and without the pool those few would run a further chunk on
    // their own, a handful of waves on an idle GPU.  ovf_pages == 0: no pool
    // (the particle ends the chunk at C records).
    char* ovf;               // ovf_pages * OVF_PAGE records
    uint32_t* ovf_first;     // per particle: its first page this chunk
    uint32_t* ovf_link;      // per page: the same particle's next page
    uint32_t* ovf_head;      // pages taken this chunk (zeroed per chunk)
    uint32_t ovf_pages;
    // trajectory-only runs (integratorType -1): no spin kernel reads the
    // records, so k_kin writes none and runs each particle to the end of the
    // call in one chunk
    uint32_t no_recs;
};

constexpr uint32_t OVF_PAGE = 256;
constexpr uint32_t OVF_NONE = 0xffffffffu;

// Address of record j of particle pid, j taken in increasing order from 0 by
// one lane (page: that lane's cursor, OVF_NONE before the first overflow
// page; it is advanced here when j opens a page).
__device__ __forceinline__ char* rec_at_seq(const PipeArgs& A, uint32_t pid, uint32_t j, uint32_t& page) {
    if (j < A.C) return A.recs + ((uint64_t)pid * A.C + j) * sizeof(SegRec);
    const uint32_t o = (j - A.C) % OVF_PAGE;
    if (o == 0) page = (j == A.C) ? A.ovf_first[pid] : A.ovf_link[page];
    return A.ovf + ((uint64_t)page * OVF_PAGE + o) * sizeof(SegRec);
}

__device__ __forceinline__ uint32_t list_len(const PipeArgs& A) { return A.n_list_dev ? *A.n_list_dev : A.n_list; }

// The physics of particle pid: its configuration's entry for a batched
// handle, the single entry for a plain one (read through the table either
// way: scalar loads where it is used instead of SGPRs held across loops).
template <bool BATCH>
__device__ __forceinline__ const Phys& lane_phys(const SimParams& P, const DevState& D, const Phys* table,
                                                 uint32_t pid) {
    if constexpr (BATCH)
        return table[(uint32_t)D.ld<uint64_t>(F_CFG, pid * 8u)];
    else
        return table[0];
}

// ---------------------------------------------------------------------------
// kinematic lane state
// ---------------------------------------------------------------------------
struct KinLane {
    uint32_t pid;
    Kin K;
    Rng rng;
    uint64_t n_steps;
    double tf;
    int32_t k;           // open interval index
    bool open, done;
    uint32_t n_events;
};

// kinematic fields (the ones k_kin reads and writes); backed up per chunk
constexpr int KIN_FIELDS[] = {F_T, F_TOLD, F_NEXTGAS, F_TF, F_PX, F_PY, F_PZ, F_POX, F_POY, F_POZ,
                              F_VX, F_VY, F_VZ, F_VOX, F_VOY, F_VOZ, F_RX, F_RY, F_RZ, F_RW, F_SPARE,
                              F_FLAGS, F_NBOUNCE, F_NCOLL, F_NSTEPS, F_KSTATE, F_NEVENTS};
constexpr int N_KIN_FIELDS = sizeof(KIN_FIELDS) / sizeof(int);

__device__ __forceinline__ void kin_load(const DevState& D, uint32_t i, KinLane& L) {
    const uint32_t o = i * 8u;
    L.pid = i;
    L.K.t = D.ld(F_T, o);
    L.K.t_old = D.ld(F_TOLD, o);
    L.K.next_gas = D.ld(F_NEXTGAS, o);
    L.K.dt = 0.0;
    L.tf = D.ld(F_TF, o);
    L.K.pos = d3{D.ld(F_PX, o), D.ld(F_PY, o), D.ld(F_PZ, o)};
    L.K.pos_old = d3{D.ld(F_POX, o), D.ld(F_POY, o), D.ld(F_POZ, o)};
    L.K.v = d3{D.ld(F_VX, o), D.ld(F_VY, o), D.ld(F_VZ, o)};
    L.K.v_old = d3{D.ld(F_VOX, o), D.ld(F_VOY, o), D.ld(F_VOZ, o)};
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
    const uint64_t ks = D.ld<uint64_t>(F_KSTATE, o);
    L.k = (int32_t)(uint32_t)ks;
    L.open = (ks >> 32) & 1;
    L.done = (ks >> 33) & 1;
    L.n_events = (uint32_t)D.ld<uint64_t>(F_NEVENTS, o);
    L.K.finished = false;
}

__device__ __forceinline__ void kin_store(const DevState& D, const KinLane& L) {
    const uint32_t o = L.pid * 8u;
    D.st(F_T, o, L.K.t);
    D.st(F_TOLD, o, L.K.t_old);
    D.st(F_NEXTGAS, o, L.K.next_gas);
    D.st(F_TF, o, L.tf);
    D.st(F_PX, o, L.K.pos.x); D.st(F_PY, o, L.K.pos.y); D.st(F_PZ, o, L.K.pos.z);
    D.st(F_POX, o, L.K.pos_old.x); D.st(F_POY, o, L.K.pos_old.y); D.st(F_POZ, o, L.K.pos_old.z);
    D.st(F_VX, o, L.K.v.x); D.st(F_VY, o, L.K.v.y); D.st(F_VZ, o, L.K.v.z);
    D.st(F_VOX, o, L.K.v_old.x); D.st(F_VOY, o, L.K.v_old.y); D.st(F_VOZ, o, L.K.v_old.z);
    D.st<uint64_t>(F_RX, o, L.rng.x);
    D.st<uint64_t>(F_RY, o, L.rng.y);
    D.st<uint64_t>(F_RZ, o, L.rng.z);
    D.st<uint64_t>(F_RW, o, L.rng.w);
    D.st(F_SPARE, o, L.rng.spare);
    D.st<uint64_t>(F_FLAGS, o, pack_flags(L.K.failure, L.K.stop, L.rng.has, L.K.coll_type, L.K.wall_hit));
    D.st<uint64_t>(F_NBOUNCE, o, L.K.n_bounce);
    D.st<uint64_t>(F_NCOLL, o, L.K.n_coll);
    D.st<uint64_t>(F_NSTEPS, o, L.n_steps);
    D.st<uint64_t>(F_KSTATE, o, (uint64_t)(uint32_t)L.k | ((uint64_t)L.open << 32) | ((uint64_t)L.done << 33));
    D.st<uint64_t>(F_NEVENTS, o, L.n_events);
}

__device__ __forceinline__ void kin_block(const RunArgs& R, const KinLane& L) {
    if (R.blocks)
        write_block_kin(R.blocks + (uint64_t)L.k * R.block_bytes, R.n, L.pid, L.K.t, L.K.pos, L.K.v, L.K.failure,
                        L.K.n_coll, L.K.n_bounce, L.n_steps);
}

__device__ __forceinline__ void put_rec(char* at, double t, const d3& p, const d3& v, uint64_t flags) {
    double* r = reinterpret_cast<double*>(at);
    r[0] = t; r[1] = p.x; r[2] = p.y; r[3] = p.z; r[4] = v.x; r[5] = v.y; r[6] = v.z;
    reinterpret_cast<uint64_t*>(r)[7] = flags;
}

// One step of the kinematic state machine: opens an interval if needed,
// runs one event of the reference's loop body, and returns the record that
// describes it (or an R_EMPTY interval close).  Shared by k_kin and k_fixup.
// It runs in three parts:
//   kin_begin  opens the interval, the collision time and the move, and names
//              the re-emission the event needs, if any;
//   run_job    that re-emission (diffuse_job / gas_job, st_device.h);
//   kin_end    sanity check, watchdog, record flags, interval close.
struct KinEvent {
    bool first;      // the event opened an interval
    bool empty;      // no event: the interval closes at once (rt/rp/rv, flags set)
    bool gas_draw;   // the gas job draws the next gas time first
    int job;         // JOB_NONE / JOB_GAS / JOB_DIFFUSE
    double Vel;      // post-flight speed (the re-emissions' input)
};

__device__ __forceinline__ void kin_begin(const SimParams& P, const Phys& F, const RunArgs& R, KinLane& L, KinEvent& E,
                                          double& rt, d3& rp, d3& rv) {
    E.first = false;
    E.empty = false;
    E.gas_draw = false;
    E.job = JOB_NONE;
    E.Vel = 0.0;
    if (!L.open) {
        // src/particle.cpp:748-752: tf = nextTOut; finished = false; tempH = h
        const double T = target_of(R, L.k);
        const double prev = (L.k == 0) ? R.t_prev : target_of(R, L.k - 1);
        bool moves = true;
        double tgt = T;
        if (R.stop_at) {
            const double e = R.stop_at[L.pid];
            if (!(prev < e)) moves = false;
            else if (e < T) tgt = e;
        }
        if (!moves || L.K.stop) {
            // the reference's loop body never runs this interval
            if (moves) L.tf = tgt;
            rt = L.K.t;
            rp = L.K.pos;
            rv = L.K.v;
            kin_block(R, L);
            L.k++;
            if (L.k >= R.n_targets) L.done = true;
            E.empty = true;
            return;
        }
        L.tf = tgt;
        L.K.finished = false;
        L.n_events = 0;
        L.open = true;
        E.first = true;
    }
    // src/particle.cpp:755-756, the re-emission left to the job
    next_collision_core(P, F, L.tf, L.K, L.rng, E.gas_draw);
    E.job = move_core(P, F, L.K, L.rng, E.Vel);
}

__device__ __forceinline__ uint64_t kin_end(const SimParams& P, const RunArgs& R, int32_t integ, KinLane& L,
                                            const KinEvent& E, double& rt, d3& rp, d3& rv) {
    if (E.empty) return R_EMPTY | R_IEND;
    // src/particle.cpp:757
    sanity(P, L.K);
    if (++L.n_events > R.max_events && L.K.failure == 0) {
        L.K.stop = true;
        L.K.failure = FAIL_WATCHDOG;
    }
    uint64_t fl = (E.first ? R_FIRST : 0) | ((uint64_t)(uint8_t)L.K.coll_type << 8) |
                  ((uint64_t)(uint8_t)L.K.wall_hit << 16);
    if (L.K.failure == 0 && integ != INT_NONE) fl |= R_INTEGRATE;
    rt = L.K.t;
    rp = L.K.pos_old;
    rv = L.K.v_old;
    if (L.K.stop) {
        fl |= R_STOP | ((uint64_t)(uint32_t)L.K.failure << 32);
        const double nan = __builtin_nan("");
        L.K.t = nan;
        L.K.pos = d3{nan, nan, nan};
        L.K.v = d3{nan, nan, nan};
    }
    L.n_steps += 1;
    if (L.K.finished || L.K.stop) {
        fl |= R_IEND;
        kin_block(R, L);
        L.k++;
        L.open = false;
        if (L.k >= R.n_targets) L.done = true;
    }
    return fl;
}

__device__ __forceinline__ void run_job(const Phys& F, KinLane& L, const KinEvent& E) {
    if (E.job == JOB_DIFFUSE) diffuse_job(L.K.v, L.K.stop, L.K.wall_hit, E.Vel, L.rng);
    else if (E.job == JOB_GAS) gas_job(L.K.v, L.K.next_gas, E.gas_draw, E.Vel, F.tc, F.sqrtKT_m, F.dist, L.rng);
}

__device__ __forceinline__ uint64_t kin_step(const SimParams& P, const Phys& F, const RunArgs& R, int32_t integ, KinLane& L,
                                             double& rt, d3& rp, d3& rv) {
    KinEvent E;
    kin_begin(P, F, R, L, E, rt, rp, rv);
    if (!E.empty) run_job(F, L, E);
    return kin_end(P, R, integ, L, E, rt, rp, rv);
}

// Resets both cursors at the start of an advance / run call.
__global__ __launch_bounds__(256) void k_begin_run(const DevState D, uint32_t n, uint32_t* list) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = i * 8u;
    D.st<uint64_t>(F_KSTATE, o, 0);
    D.st<uint64_t>(F_SSTATE, o, 0);
    D.st<uint64_t>(F_NEVENTS, o, 0);
    // kinematics and spin are in step between calls: the first segment of
    // this call starts at the particle's current time
    D.st(F_TSPIN, o, D.ld(F_T, o));
    list[i] = i;
}

// waiting lanes that start a wave's diffuse re-emissions (He-3, 1M, 4 x 0.5 s:
// k_kin 105 ms at 4, 106 at 8, 114 at 16, 108 without the wait); a particle
// whose walls are mostly diffuse takes its re-emission at once
constexpr int DIFFUSE_BATCH = 4;
template <bool BATCH>
__global__ __launch_bounds__(256) void k_kin(const SimParams P, const DevState D, const PipeArgs A) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt[2] = {0, 0};   // events, trips
    if (li < list_len(A)) {
        KinLane L;
        kin_load(D, A.list[li], L);
        const Phys& F = lane_phys<BATCH>(P, D, A.phys, L.pid);
        // back up the kinematic state for a possible spin-failure replay
        const uint32_t o = L.pid * 8u;
        const DevState B{A.backup, D.stride};
#pragma unroll
        for (int f = 0; f < N_KIN_FIELDS; ++f) B.st<uint64_t>(KIN_FIELDS[f], o, D.ld<uint64_t>(KIN_FIELDS[f], o));
        uint32_t j = 0, lim = A.no_recs ? 0xffffffffu : A.C, page = OVF_NONE, rpage = OVF_NONE;
        const double t0 = L.K.t;
        KinEvent E;
        bool pend = false;   // the event's diffuse re-emission waits for more lanes
        const int batch = F.diffuse < 0.5 ? DIFFUSE_BATCH : 1;
        double rt;
        d3 rp, rv;
        while (!L.done) {
            if (!pend && j == lim) {
                // Chunk depth used up: continue on an overflow page if one is
                // left and the particle is near its end -- its remaining time
                // to the last output under twice what this chunk has covered
                // -- or the run has a single output time (every particle
                // then needs the same interval; the pages keep its heavy
                // tail out of a chunk of its own).  Pages taken by particles
                // that need several chunks anyway would only lengthen this
                // chunk for their lanes (He-3, 1M, 20 x 0.5 s at the default
                // depth: 1,432 ms with this rule, 1,455 ms taking pages for
                // any particle, 1,538 ms with four times the covered time).
                if (A.ovf_pages == 0) break;
                double t_last = target_of(A.R, A.R.n_targets - 1);
                if (A.R.stop_at) t_last = fmin(t_last, A.R.stop_at[L.pid]);
                if (A.R.n_targets > 1 && !(t_last - L.K.t <= 2.0 * (L.K.t - t0))) break;
                const uint32_t p = atomicAdd(A.ovf_head, 1u);
                if (p >= A.ovf_pages) break;
                if (page == OVF_NONE) A.ovf_first[L.pid] = p;
                else A.ovf_link[page] = p;
                page = p;
                lim += OVF_PAGE;
            }
            if (!pend) {
                kin_begin(P, F, A.R, L, E, rt, rp, rv);
                if (!E.empty) {
                    if (E.job == JOB_GAS)
                        gas_job(L.K.v, L.K.next_gas, E.gas_draw, E.Vel, F.tc, F.sqrtKT_m, F.dist, L.rng);
                    pend = E.job == JOB_DIFFUSE;
                }
            }
            // The diffuse re-emission (acos and two sincos; about a sixth of
            // this kernel's time, for ~3% of He-3 events) would run on almost
            // every trip for a lane or two.  It runs once `batch` lanes of the
            // wave wait for it, or when no running lane has other work;
            // a waiting lane idles, its random stream untouched meanwhile, so
            // every lane's events and draws are those of kin_step.
            {
                const uint64_t waiting = __ballot(pend), running = __ballot(1);
                if (pend && (__popcll(waiting) >= batch || waiting == running)) {
                    diffuse_job(L.K.v, L.K.stop, L.K.wall_hit, E.Vel, L.rng);
                    pend = false;
                }
            }
            if (pend) continue;
            const uint64_t fl = kin_end(P, A.R, A.integ, L, E, rt, rp, rv);
            if (!A.no_recs) put_rec(rec_at_seq(A, L.pid, j, rpage), rt, rp, rv, fl);
            ++j;
            cnt[0] += (fl & R_EMPTY) ? 0 : 1;
        }
        cnt[1] = j;
        A.seg_count[L.pid] = j;
        kin_store(D, L);
        if (!L.done) A.next_list[atomicAdd(A.next_count, 1u)] = L.pid;
    }
    flush_counters(A.R.work + 4, cnt);
}

// ---------------------------------------------------------------------------
// spin lane
// ---------------------------------------------------------------------------
struct SpinLane {
    uint32_t pid, j, nrec;
    uint32_t page;       // overflow-page cursor of the record reads
    uint32_t cfg;        // batched handles: the particle's configuration
    int32_t k;           // interval index of the next snapshot block
    bool stopped;
    double h, tempH, t_spin;
    d3 S;
    uint32_t acc, rej, seg;
};

template <bool BATCH>
__device__ __forceinline__ void spin_load(const DevState& D, const PipeArgs& A, uint32_t pid, SpinLane& L) {
    const uint32_t o = pid * 8u;
    L.pid = pid;
    L.cfg = BATCH ? (uint32_t)D.ld<uint64_t>(F_CFG, o) : 0u;
    L.j = 0;
    L.page = OVF_NONE;
    L.nrec = A.seg_count[pid];
    const uint64_t ss = D.ld<uint64_t>(F_SSTATE, o);
    L.k = (int32_t)(uint32_t)ss;
    L.stopped = (ss >> 33) & 1;
    L.h = D.ld(F_H, o);
    L.tempH = D.ld(F_TEMPH, o);
    L.t_spin = D.ld(F_TSPIN, o);
    L.S = d3{D.ld(F_SX, o), D.ld(F_SY, o), D.ld(F_SZ, o)};
    L.acc = L.rej = L.seg = 0;
    // complete these loads here (a particle switch is rare): left in flight,
    // they put conservative waits into the trial loop, which then also wait
    // for the other lanes' snapshot and state stores
    asm volatile("" ::"v"(L.h), "v"(L.tempH), "v"(L.t_spin), "v"(L.S.x), "v"(L.S.y), "v"(L.S.z), "v"(L.nrec),
                 "v"(ss), "v"(L.cfg));
}

__device__ __forceinline__ void spin_store(const DevState& D, const SpinLane& L) {
    const uint32_t o = L.pid * 8u;
    D.st<uint64_t>(F_SSTATE, o, (uint64_t)(uint32_t)L.k | ((uint64_t)L.stopped << 33));
    D.st(F_H, o, L.h);
    D.st(F_TEMPH, o, L.tempH);
    D.st(F_TSPIN, o, L.t_spin);
    D.st(F_SX, o, L.S.x); D.st(F_SY, o, L.S.y); D.st(F_SZ, o, L.S.z);
    D.st<uint64_t>(F_ACC, o, D.ld<uint64_t>(F_ACC, o) + L.acc);
    D.st<uint64_t>(F_REJ, o, D.ld<uint64_t>(F_REJ, o) + L.rej);
    D.st<uint64_t>(F_SEG, o, D.ld<uint64_t>(F_SEG, o) + L.seg);
}

__device__ __forceinline__ SegRec load_rec(const char* at) {
    const double* r = reinterpret_cast<const double*>(at);
    SegRec s;
    s.t = r[0]; s.px = r[1]; s.py = r[2]; s.pz = r[3]; s.vx = r[4]; s.vy = r[5]; s.vz = r[6];
    s.flags = reinterpret_cast<const uint64_t*>(r)[7];
    return s;
}
// record j of particle pid in the fixed layout (no overflow pages)
__device__ __forceinline__ SegRec get_rec(const char* recs, uint32_t C, uint32_t pid, uint32_t j) {
    return load_rec(recs + ((uint64_t)pid * C + j) * sizeof(SegRec));
}

__device__ __forceinline__ void spin_block(const RunArgs& R, const SpinLane& L) {
    if (R.blocks) write_block_spin(R.blocks + (uint64_t)L.k * R.block_bytes, R.n, L.pid, L.S);
}

template <int VAR, int INTEG, bool BATCH, bool PULSE>
__device__ __forceinline__ void spin_body(const SimParams& P, const DevState& D, const PipeArgs& A) {
    const RunArgs& R = A.R;
    SpinLane L;
    // physics of the lane's current particle
    // (a plain handle's single entry is read through the same table: uniform
    // scalar loads at each segment set-up instead of ~50 SGPRs held across
    // the loop)
    auto phys = [&]() -> const Phys& {
        if constexpr (BATCH)
            return A.phys[L.cfg];
        else
            return A.phys[0];
    };
    DopState dop;
    RkState rk;
    // end time of the segment being integrated (the record's time)
    auto seg_end = [&]() -> double {
        if constexpr (INTEG == INT_DOP853)
            return dop.xf;
        else
            return rk.tf;
    };
    Seg seg;
    Poly poly;
    uint32_t cur_flags = 0;   // low flag bits of the record being integrated (IEND, STOP, coll_type, wall_hit)
    unsigned long long cnt[4] = {0, 0, 0, 0};   // accepted, rejected, segments, trips
    enum { S_DONE = 0, S_NEXT = 1, S_INT = 2 };
    int phase = S_DONE;
    const uint32_t n_list = list_len(A);
    {
        const uint32_t q = acquire_index(R.queue);
        if (q < n_list) {
            spin_load<BATCH>(D, A, A.list[q], L);
            phase = S_NEXT;
        }
    }
    while (phase != S_DONE) {
        ++cnt[3];
        if (phase == S_NEXT) {
            // consume records until one needs integration (or the chunk ends)
            while (phase == S_NEXT) {
                if (L.j >= L.nrec) {
                    cnt[0] += L.acc; cnt[1] += L.rej; cnt[2] += L.seg;
                    spin_store(D, L);
                    phase = S_DONE;
                    const uint32_t q = acquire_index(R.queue);
                    if (q < n_list) {
                        spin_load<BATCH>(D, A, A.list[q], L);
                        phase = S_NEXT;
                        continue;
                    }
                    break;
                }
                const SegRec rec = load_rec(rec_at_seq(A, L.pid, L.j, L.page));
                cur_flags = (uint32_t)rec.flags;
                if (rec.flags & R_FIRST) L.tempH = L.h;
                if ((rec.flags & R_INTEGRATE) && !L.stopped) {
                    seg.t0 = L.t_spin;
                    seg.p0 = d3{rec.px, rec.py, rec.pz};
                    seg.v0 = d3{rec.vx, rec.vy, rec.vz};
                    if (VAR == VAR_FAST) poly = make_poly(P, phys(), seg);
                    if (INTEG == INT_DOP853) {
                        if (VAR == VAR_EXACT) dop_begin_exact(P, phys(), dop, L.S, seg, rec.t);
                        else dop_begin_fast(P, phys(), dop, L.S, seg, poly, rec.t);
                    } else {
                        rk_begin(rk, seg, rec.t, L.tempH);
                    }
                    L.seg++;
                    if (R.trace && L.pid == R.trace_pid)
                        ST_TRACE(R, 1, seg.t0, rec.t, seg.p0, seg.v0, L.S, L.tempH, L.acc, L.rej,
                                     (char)(rec.flags >> 8), (char)(rec.flags >> 16), 0, 0, 0);
                    phase = S_INT;
                } else {
                    // no integration: sanity failure, trajectory-only run, or an empty close
                    if (rec.flags & R_STOP) {
                        const double nan = __builtin_nan("");
                        L.S = d3{nan, nan, nan};
                    }
                    if (rec.flags & R_IEND) {
                        spin_block(R, L);
                        L.k++;
                    }
                    if (!(rec.flags & R_EMPTY)) L.t_spin = rec.t;
                    L.j++;
                }
            }
        }
        if (phase == S_INT) {
            int st;
            if constexpr (INTEG == INT_DOP853) {
                if (VAR == VAR_EXACT) st = dop_trial_exact(P, phys(), dop, L.S, L.tempH, seg, L.acc, L.rej);
                else st = dop_trial_fast<PULSE>(P, phys(), dop, L.S, L.tempH, seg, poly, L.acc, L.rej);
            } else {
                st = rk_any_trial<VAR, INTEG>(P, phys(), rk, L.S, L.tempH, seg, poly, L.acc, L.rej);
            }
            if (st != 0) {
                if (R.trace && L.pid == R.trace_pid)
                    ST_TRACE(R, 2, L.t_spin, seg_end(), d3{0, 0, 0}, d3{0, 0, 0}, L.S, L.tempH, L.acc, L.rej,
                             (char)(cur_flags >> 8), (char)(cur_flags >> 16), st, 0, 0);
                // src/particle.cpp:593-612
                if (P.keep) L.h = L.tempH;
                struct {
                    uint32_t flags;
                } rec{cur_flags};
                L.t_spin = seg_end();
                if (st < 0) {
                    // the reference stops the particle here; the kinematics ran
                    // ahead, so k_fixup replays it up to this record
                    const uint32_t slot = atomicAdd(A.fail_count, 1u);
                    if (slot < A.fail_cap) {
                        A.fail_list[3 * slot + 0] = L.pid;
                        A.fail_list[3 * slot + 1] = L.j;
                        A.fail_list[3 * slot + 2] = (uint32_t)st;
                    }
                    L.stopped = true;
                }
                if (L.stopped || (rec.flags & R_STOP)) {
                    const double nan = __builtin_nan("");
                    L.S = d3{nan, nan, nan};
                }
                if ((rec.flags & R_IEND) || st < 0) {
                    // (a failing particle's later blocks are written by k_fixup)
                    spin_block(R, L);
                    L.k++;
                }
                L.j++;
                phase = S_NEXT;
            }
        }
    }
    flush_counters(R.work, cnt);
}

template <int VAR, int INTEG, bool BATCH, bool PULSE>
__global__ __launch_bounds__(256) void k_spin(const SimParams P, const DevState D, const PipeArgs A) {
    spin_body<VAR, INTEG, BATCH, PULSE>(P, D, A);
}
// The same kernel held to 168 VGPRs (three waves per SIMD); the fast DOP853
// trial without the pulse term fits there, and the third wave hides the FP64
// dependency latency the second one leaves exposed.
template <int VAR, int INTEG, bool BATCH, bool PULSE>
__global__ __launch_bounds__(256, 3) void k_spin_o3(const SimParams P, const DevState D, const PipeArgs A) {
    spin_body<VAR, INTEG, BATCH, PULSE>(P, D, A);
}
// ... and held to 128 (four waves per SIMD), for the quaternion trials
template <int VAR, int INTEG, bool BATCH, bool PULSE>
__global__ __launch_bounds__(256, 4) void k_spin_o4(const SimParams P, const DevState D, const PipeArgs A) {
    spin_body<VAR, INTEG, BATCH, PULSE>(P, D, A);
}

// Rare path: a spin-integration failure at record j_fail of this chunk.
// Rewind the kinematics to the chunk start, replay exactly j_fail+1 records
// (the failing segment's event included), then apply the reference's stop:
// failureState = status, t/pos/v/S NaN, and every later snapshot block shows
// the stopped particle.
// One failing particle per lane, grid-stride; n_fail == FIXUP_FROM_DEVICE
// reads the count the spin kernel left in *A.fail_count (graph-captured
// chunks, where the host does not see it).
constexpr uint32_t FIXUP_FROM_DEVICE = 0xffffffffu;
template <bool BATCH>
__device__ __forceinline__ void fixup_one(const SimParams& P, const DevState& D, const PipeArgs& A, uint32_t f);

template <bool BATCH>
__global__ __launch_bounds__(64) void k_fixup(const SimParams P, const DevState D, const PipeArgs A,
                                              uint32_t n_fail) {
    const uint32_t nf = n_fail == FIXUP_FROM_DEVICE ? min(*A.fail_count, A.fail_cap) : n_fail;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x)
        fixup_one<BATCH>(P, D, A, f);
}

template <bool BATCH>
__device__ __forceinline__ void fixup_one(const SimParams& P, const DevState& D, const PipeArgs& A, uint32_t f) {
    const uint32_t pid = A.fail_list[3 * f + 0];
    const uint32_t jf = A.fail_list[3 * f + 1];
    const int32_t status = (int32_t)A.fail_list[3 * f + 2];
    const uint32_t o = pid * 8u;
    const DevState B{A.backup, D.stride};
    for (int q = 0; q < N_KIN_FIELDS; ++q) D.st<uint64_t>(KIN_FIELDS[q], o, B.ld<uint64_t>(KIN_FIELDS[q], o));
    KinLane L;
    kin_load(D, pid, L);
    const Phys& F = lane_phys<BATCH>(P, D, A.phys, pid);
    RunArgs R = A.R;
    R.blocks = nullptr;   // blocks are rewritten below
    int32_t k_at = L.k;
    for (uint32_t j = 0; j <= jf; ++j) {
        k_at = L.k;
        double rt;
        d3 rp, rv;
        (void)kin_step(P, F, R, A.integ, L, rt, rp, rv);
    }
    // the failing segment belongs to interval k_at; the stop happens there
    const double nan = __builtin_nan("");
    L.K.stop = true;
    L.K.failure = status;
    L.K.t = nan;
    L.K.pos = d3{nan, nan, nan};
    L.K.v = d3{nan, nan, nan};
    if (A.R.blocks) {
        const d3 Sn{nan, nan, nan};
        for (int32_t k = k_at; k < A.R.n_targets; ++k)
            write_record(A.R.blocks + (uint64_t)k * A.R.block_bytes, A.R.n, pid, L.K.t, L.K.pos, L.K.v, Sn,
                         L.K.failure, L.K.n_coll, L.K.n_bounce, L.n_steps);
    }
    L.k = A.R.n_targets;
    L.open = false;
    L.done = true;
    kin_store(D, L);
    D.st<uint64_t>(F_SSTATE, o, (uint64_t)(uint32_t)A.R.n_targets | (1ull << 33));
}

}  // namespace st
