// st_engine.hip -- gfx950 kernels of the spin-tracking hot path and the C ABI
// declared in include/spintrack.h.
//
// Design (one particle per lane, wave64, FP64 on the vector ALUs):
//  * Particle state lives in HBM as structure-of-arrays (one 8-byte array per
//    field), loaded and stored coalesced.
//  * Default engine: the two-kernel pipeline (st_pipeline.h).  k_kin runs the
//    branchy event loop and stages one 64-byte record per trajectory segment
//    in HBM; k_spin runs one DOP853/RK45 trial per loop trip over those
//    records, refilling lanes from an atomic work queue; runs are cut into
//    chunks whose record buffer fits a configurable HBM budget.
//  * Alternative engine: the fused single-kernel loop (st_fused.h).
//  * Snapshot blocks are written on the GPU, per lane, as each particle
//    crosses an output time, already in the reference's file layout.
//  * Tableau coefficients are compile-time constants (SGPR literals); run
//    parameters are one uniform kernel-argument block.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/spintrack.h"
#include "st_adjoint.h"
#include "st_analysis.h"
#include "st_floquet.h"
#include "st_fused.h"
#include "st_pipeline.h"
#include "st_spectrum.h"

static_assert(sizeof(st_options) == 440, "st_options must match the reference's 440-byte options");
static_assert(st::FAIL_WATCHDOG == ST_FAIL_WATCHDOG, "watchdog failure code");
static_assert(offsetof(st_options, nmax) == 360, "st_options layout");
static_assert(offsetof(st_options, dist) == 384, "st_options layout");
static_assert(offsetof(st_options, noiseAmplitudes) == 392, "st_options layout");
static_assert(sizeof(st_rng) == 48, "st_rng must match the reference's 48-byte rngState");

namespace st {

// initParticlesCPU (src/particle.cpp:638-710), one lane per particle; local
// slot j is global particle first + j*stride.  A batched handle lays its
// configurations out back to back (cfg_off[c] .. cfg_off[c+1]); slot j of
// configuration c is seeded as particle j of a plain run with that
// configuration's seed, so each configuration reproduces its own plain run.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_init(const SimParams P, const DevState D, uint32_t n, uint32_t seed,
                                              uint32_t first, uint32_t stride, double t0, double tf0,
                                              const Phys* phys, const uint32_t* cfg_off, const uint32_t* cfg_seed,
                                              uint32_t n_cfg) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t cfg = 0, local = j;
    if constexpr (BATCH) {
        uint32_t lo = 0, hi = n_cfg;   // last c with cfg_off[c] <= j
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (cfg_off[mid] <= j) lo = mid;
            else hi = mid;
        }
        cfg = lo;
        local = j - cfg_off[cfg];
        seed = cfg_seed[cfg];
    }
    const Phys& F = BATCH ? phys[cfg] : P.ph;
    const uint32_t ipart = first + local * stride;
    Rng rng;
    seed_rng(rng, seed, ipart);
    d3 p;
    p.x = uni(rng) * P.Lx - P.Lx / 2.0;
    p.y = uni(rng) * P.Ly - P.Ly / 2.0;
    p.z = uni(rng) * P.Lz - P.Lz / 2.0;
    double ng;
    if (P.gas_coll)
        ng = expo_draw(rng, F.tc);
    else
        ng = DBL_MAX_D;
    d3 v{0.0, 0.0, 0.0};
    if (F.dist == 'C' || F.dist == 'M') {
        double n0, n1, n2;
        normal3(rng, n0, n1, n2);
        if (F.dist == 'C') {
            const double V = F.V;
            const double vn = len3(d3{n0, n1, n2});
            v = d3{(V * n0) / vn, (V * n1) / vn, (V * n2) / vn};
        } else {
            v = d3{n0 * F.sqrtKT_m, n1 * F.sqrtKT_m, n2 * F.sqrtKT_m};
        }
    }
    const uint32_t o = j * 8u;
    D.st(F_PX, o, p.x); D.st(F_PY, o, p.y); D.st(F_PZ, o, p.z);
    D.st(F_POX, o, p.x); D.st(F_POY, o, p.y); D.st(F_POZ, o, p.z);
    D.st(F_NEXTGAS, o, ng);
    D.st(F_VX, o, v.x); D.st(F_VY, o, v.y); D.st(F_VZ, o, v.z);
    D.st(F_VOX, o, v.x); D.st(F_VOY, o, v.y); D.st(F_VOZ, o, v.z);
    D.st<uint64_t>(F_RX, o, rng.x);
    D.st<uint64_t>(F_RY, o, rng.y);
    D.st<uint64_t>(F_RZ, o, rng.z);
    D.st<uint64_t>(F_RW, o, rng.w);
    D.st(F_SPARE, o, rng.spare);
    D.st(F_SX, o, F.yix); D.st(F_SY, o, F.yiy); D.st(F_SZ, o, F.yiz);
    D.st(F_T, o, t0);
    D.st(F_TOLD, o, t0);
    D.st(F_TF, o, tf0);
    D.st(F_H, o, P.h0);
    D.st(F_TEMPH, o, P.h0);
    D.st(F_TSPIN, o, t0);
    // coll_type 'W', wall_hit 0 (uninitialised in the reference), no failure
    D.st<uint64_t>(F_FLAGS, o, pack_flags(0, false, rng.has, 'W', 0));
    D.st<uint64_t>(F_NBOUNCE, o, 0);
    D.st<uint64_t>(F_NCOLL, o, 0);
    D.st<uint64_t>(F_NSTEPS, o, 0);
    D.st<uint64_t>(F_PARTID, o, ipart);
    D.st<uint64_t>(F_ACC, o, 0);
    D.st<uint64_t>(F_REJ, o, 0);
    D.st<uint64_t>(F_SEG, o, 0);
    D.st<uint64_t>(F_KSTATE, o, 0);
    D.st<uint64_t>(F_SSTATE, o, 0);
    D.st<uint64_t>(F_NEVENTS, o, 0);
    D.st<uint64_t>(F_CFG, o, cfg);
}

// Graph-captured chunk loop: the counters a chunk uses live in small[]
// ([0] spin queue head, [1] next-list length, [2] failures, [4] this chunk's
// list length).  Single-thread kernels stand in for host memsets, which a
// captured graph must not hold.
__global__ void k_graph_start(uint32_t* small, uint32_t n) {
    small[0] = 0;
    small[1] = n;   // the "next list" of chunk -1: every particle, in list 0
    small[2] = 0;
    small[4] = 0;
}
// the list the previous chunk left becomes this chunk's; counters restart
__global__ void k_chunk_head(uint32_t* small) {
    small[4] = small[1];
    small[0] = 0;
    small[1] = 0;
    small[2] = 0;
}

// Current state -> one snapshot block (reference file layout).  A workgroup
// owns 256 consecutive particles.  t, failure and the counters are one
// element per lane and go straight out (consecutive lanes, consecutive
// addresses).  The coordinate fields are 24-byte {x, y, z} records: each is
// staged in LDS as the byte image of the group's 6 KiB slice and written back
// as contiguous 8-byte stores (2 KiB per store instruction across the group).
__global__ __launch_bounds__(256) void k_snapshot(const DevState D, uint32_t n, char* B) {
    __shared__ double stage[3 * 256];
    const uint32_t tid = threadIdx.x;
    const uint32_t base = blockIdx.x * 256u;
    const uint32_t i = base + tid;
    const uint32_t cnt = min(256u, n - base);
    const bool live = i < n;
    const uint32_t o = i * 8u;
    const uint64_t N = n;
    if (live) {
        reinterpret_cast<double*>(B)[i] = D.ld(F_T, o);
        reinterpret_cast<int32_t*>(B + 80 * N)[i] = (int32_t)(uint32_t)D.ld<uint64_t>(F_FLAGS, o);
        put_u64(B + 84 * N + 8 * (uint64_t)i, D.ld<uint64_t>(F_NCOLL, o));
        put_u64(B + 92 * N + 8 * (uint64_t)i, D.ld<uint64_t>(F_NBOUNCE, o));
        put_u64(B + 100 * N + 8 * (uint64_t)i, D.ld<uint64_t>(F_NSTEPS, o));
    }
    const int fx[3] = {F_PX, F_VX, F_SX};
    const uint64_t off[3] = {8 * N, 32 * N, 56 * N};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        if (live) {
            stage[3 * tid + 0] = D.ld(fx[f] + 0, o);
            stage[3 * tid + 1] = D.ld(fx[f] + 1, o);
            stage[3 * tid + 2] = D.ld(fx[f] + 2, o);
        }
        __syncthreads();
        double* dst = reinterpret_cast<double*>(B + off[f]) + 3 * (uint64_t)base;
        for (uint32_t k = tid; k < 3 * cnt; k += 256) dst[k] = stage[k];
        __syncthreads();
    }
}

}  // namespace st

// ===========================================================================
// C ABI
// ===========================================================================
using namespace st;

namespace {
thread_local std::string g_create_error;
constexpr int N_WORK = 8;   // spin: accepted, rejected, segments, trips; kin: events, trips
}  // namespace

struct st_handle {
    st_options opt;
    st_ext ext;
    SimParams P;
    DevState D;
    uint32_t n = 0;
    void* state = nullptr;
    size_t state_bytes = 0;
    double* d_targets = nullptr;
    int32_t targets_cap = 0;
    double* d_stop_at = nullptr;
    bool have_stop_at = false;
    unsigned long long* d_work = nullptr;
    uint32_t* d_small = nullptr;   // [0] queue, [1] next_count, [2] fail_count, [3] overflow pages taken
    int device = -1;               // the device current at st_create
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool initialized = false;
    double last_target = 0.0;
    int grid_fused = 0, grid_spin = 0;
    // pipeline buffers
    char* d_recs = nullptr;
    uint32_t rec_cap = 0;          // C
    bool no_recs = false;          // trajectory-only: k_kin writes no records
    char* d_ovf = nullptr;         // overflow pages (plain pipeline runs)
    uint32_t* d_ovf_first = nullptr;
    uint32_t* d_ovf_link = nullptr;
    uint32_t ovf_pages = 0;
    uint32_t* d_list[2] = {nullptr, nullptr};
    uint32_t* d_seg_count = nullptr;
    char* d_backup = nullptr;
    uint32_t* d_fail = nullptr;
    // the adjoint's replay window (several checkpoint chunks of records)
    char* d_adj_recs = nullptr;
    size_t adj_recs_bytes = 0;
    // segment trace of one particle
    double* d_trace = nullptr;
    uint32_t* d_trace_count = nullptr;
    uint32_t trace_cap = 0;
    int64_t trace_pid = -1;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // per-kernel timing of the pipeline (st_kernel_times)
    bool ktime_on = false;
    hipEvent_t kev[4] = {nullptr, nullptr, nullptr, nullptr};
    double kms[2] = {0.0, 0.0};     // k_kin, k_spin milliseconds
    uint64_t klaunch[2] = {0, 0};   // launches
    // batched handles: configurations back to back in particle order
    bool batch = false;
    bool any_pulse = false;          // some configuration has a pulse term
    uint32_t n_cfg = 1;
    std::vector<uint32_t> cfg_off;   // n_cfg + 1 offsets
    Phys* d_phys = nullptr;
    uint32_t* d_cfg_off = nullptr;
    uint32_t* d_cfg_seed = nullptr;
    // adjoint checkpoints of the last pipeline run: kinematic state at each
    // chunk start and that chunk's active list
    struct Ckpt {
        char* state;
        uint32_t* list;
        uint32_t n_list;
    };
    bool ckpt_on = false;
    std::vector<Ckpt> ckpts;
    RunArgs ckpt_run;               // the run the checkpoints belong to
    double* d_ckpt_targets = nullptr;
    // adjoint work buffers
    char* d_adj = nullptr;          // A_COUNT fields
    char* d_replay = nullptr;       // kinematic state of the replayed chunk
    char* d_scratch = nullptr;      // next list, counters, failure list of replays
    // spectral path (integratorType 6): per-particle aggregators, the run's
    // summed covariance spectrum and the Floquet decomposition of the drive
    SpecArgs spec;                  // aggregator pointers and filter constants
    double* d_spec = nullptr;       // SPEC_ENTRIES doubles
    unsigned long long* d_spec_n = nullptr;
    double* d_spec_part = nullptr;  // SPEC_RED_BLOCKS * (SPEC_ENTRIES + 1)
    floquet::Diag fd;
    int grid_spec = 0;
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

#define HIPCHK(h, call)                                                                 \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail((h), ST_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static Phys derive_phys(const st_options& o, const st_ext& x) {
    Phys F;
    memset(&F, 0, sizeof F);
    F.B0x = o.B0.x; F.B0y = o.B0.y; F.B0z = o.B0.z;
    F.Ex = o.E.x; F.Ey = o.E.y; F.Ez = o.E.z;
    F.G[0] = o.Gx.x; F.G[1] = o.Gx.y; F.G[2] = o.Gx.z;
    F.G[3] = o.Gy.x; F.G[4] = o.Gy.y; F.G[5] = o.Gy.z;
    F.G[6] = o.Gz.x; F.G[7] = o.Gz.y; F.G[8] = o.Gz.z;
    F.gamma = o.gamma;
    F.pa = o.a;
    F.pw = o.w;
    F.pulse_axis = x.pulse_axis;
    const double b0ax = x.pulse_axis == 0 ? o.B0.x : (x.pulse_axis == 1 ? o.B0.y : o.B0.z);
    // a*cos(w*t) with a == 0 is a signed zero; added to B0[axis] it changes
    // nothing unless B0[axis] is itself -0.0 (or cos is not finite).
    // (a pulse hook of another form is evaluated always, as the reference's is)
    F.pulse_on = (!ST_FIELDS_PULSE_COS_AXIS || o.a != 0.0 || signbit(b0ax) || !isfinite(o.w)) ? 1 : 0;
    F.tc = o.tc;
    F.sqrtKT_m = o.sqrtKT_m;
    F.diffuse = o.diffuse;
    F.diffuse_on = (o.diffuse > (double)1.17549435082228750797e-38) ? 1 : 0;
    F.V = o.V;
    F.yix = o.yi.x; F.yiy = o.yi.y; F.yiz = o.yi.z;
    F.dist = (int32_t)(unsigned char)o.dist;
    set_poly_consts(F, o.gravity);
    return F;
}

static SimParams derive_params(const st_options& o, const st_ext& x) {
    SimParams P;
    memset(&P, 0, sizeof P);
    P.Lx = o.L.x; P.Ly = o.L.y; P.Lz = o.L.z;
    P.rtol = o.rtol; P.atol = o.atol; P.beta = o.beta; P.safe = o.safe;
    P.fac1 = o.fac1; P.fac2 = o.fac2; P.hmax = o.hmax; P.hmin = o.hmin; P.h0 = o.h;
    P.expo1 = 1.0 / 8.0 - o.beta * 0.2;
    P.facc1 = 1.0 / o.fac1;
    P.facc2 = 1.0 / o.fac2;
    P.pow_beta_on = (o.beta != 0.0) ? 1 : 0;
    P.maxPosStep = o.maxPosStep;
    P.nmax = o.nmax;
    P.gravity = o.gravity ? 1 : 0;
    P.gas_coll = o.gas_coll ? 1 : 0;
    P.fixed = o.fixedStepSize ? 1 : 0;
    P.keep = o.keepStepSize ? 1 : 0;
    P.ph = derive_phys(o, x);
    return P;
}

static uint64_t field_stride(uint32_t n) { return ((uint64_t)8 * n + 255) & ~(uint64_t)255; }

// ---- kernel selection ------------------------------------------------------
template <int VAR, int INTEG, bool PERSIST, bool PULSE = true>
static hipError_t launch_fused(st_handle* h, dim3 grid, dim3 block, const RunArgs& R) {
    hipLaunchKernelGGL((k_engine<VAR, INTEG, PERSIST, PULSE>), grid, block, 0, h->stream, h->P, h->D, R);
    return hipGetLastError();
}
using FusedFn = hipError_t (*)(st_handle*, dim3, dim3, const RunArgs&);
template <bool PERSIST>
static FusedFn pick_fused(int variant, int integ, bool pulse) {
    if (variant == ST_VARIANT_FAST) {
        if (integ == 0)
            return pulse ? launch_fused<VAR_FAST, INT_DOP853, PERSIST, true>
                         : launch_fused<VAR_FAST, INT_DOP853, PERSIST, false>;
        if (integ == 1) return launch_fused<VAR_FAST, INT_RK45, PERSIST>;
        return launch_fused<VAR_FAST, INT_NONE, PERSIST>;
    }
    if (integ == 0) return launch_fused<VAR_EXACT, INT_DOP853, PERSIST>;
    if (integ == 1) return launch_fused<VAR_EXACT, INT_RK45, PERSIST>;
    return launch_fused<VAR_EXACT, INT_NONE, PERSIST>;
}
static const void* fused_symbol(int variant, int integ, bool pulse) {
    if (variant == ST_VARIANT_FAST)
        return integ == 0 ? (pulse ? (const void*)&k_engine<VAR_FAST, INT_DOP853, true, true>
                                   : (const void*)&k_engine<VAR_FAST, INT_DOP853, true, false>)
                          : (const void*)&k_engine<VAR_FAST, INT_RK45, true>;
    return integ == 0 ? (const void*)&k_engine<VAR_EXACT, INT_DOP853, true>
                      : (const void*)&k_engine<VAR_EXACT, INT_RK45, true>;
}

// Spin kernels.  The fast DOP853 kernel has a form without the pulse term
// (PULSE = false: fewer live registers) for handles whose configurations
// have no pulse; the other kernels evaluate the pulse switch at run time.
// The fast DOP853 kernels run at occupancy 3 (the pulse form spilling ~90 B
// per lane to do so), except the batched pulse form, whose per-lane pulse
// fields make the spills cost more than the third wave gains; so does the
// fast MagnusCFET kernel, which sits at the 168-register line.  The fast
// quaternion kernels of plain handles run at occupancy 4 (128 registers,
// about 28 B of spills per lane): the fourth wave covers more of the
// dependent rotation chains than the spills cost.
template <int VAR, int INTEG, bool BATCH, bool PULSE>
constexpr bool use_o3() {
    return VAR == VAR_FAST && (INTEG == INT_DOP853 || INTEG == INT_MAGNUS);
}
template <int VAR, int INTEG, bool BATCH, bool PULSE>
constexpr bool use_o4() {
    return VAR == VAR_FAST && !BATCH && (INTEG == INT_RK45Q || INTEG == INT_RKF45Q);
}
template <int VAR, int INTEG, bool BATCH, bool PULSE>
static hipError_t launch_spin(st_handle* h, dim3 grid, const PipeArgs& A) {
    if constexpr (use_o4<VAR, INTEG, BATCH, PULSE>())
        hipLaunchKernelGGL((k_spin_o4<VAR, INTEG, BATCH, PULSE>), grid, dim3(256), 0, h->stream, h->P, h->D, A);
    else if constexpr (use_o3<VAR, INTEG, BATCH, PULSE>())
        hipLaunchKernelGGL((k_spin_o3<VAR, INTEG, BATCH, PULSE>), grid, dim3(256), 0, h->stream, h->P, h->D, A);
    else
        hipLaunchKernelGGL((k_spin<VAR, INTEG, BATCH, PULSE>), grid, dim3(256), 0, h->stream, h->P, h->D, A);
    return hipGetLastError();
}
using SpinFn = hipError_t (*)(st_handle*, dim3, const PipeArgs&);
template <int VAR, bool BATCH>
static SpinFn pick_spin_v(int integ, bool pulse) {
    switch (integ) {
    case INT_RK45: return launch_spin<VAR, INT_RK45, BATCH, true>;
    case INT_MAGNUS: return launch_spin<VAR, INT_MAGNUS, BATCH, true>;
    case INT_RK45Q: return launch_spin<VAR, INT_RK45Q, BATCH, true>;
    case INT_RKF45: return launch_spin<VAR, INT_RKF45, BATCH, true>;
    case INT_RKF45Q: return launch_spin<VAR, INT_RKF45Q, BATCH, true>;
    default:
        if constexpr (VAR == VAR_FAST)
            return pulse ? launch_spin<VAR, INT_DOP853, BATCH, true> : launch_spin<VAR, INT_DOP853, BATCH, false>;
        else
            return launch_spin<VAR, INT_DOP853, BATCH, true>;
    }
}
static SpinFn pick_spin(int variant, int integ, bool batch, bool pulse) {
    if (variant == ST_VARIANT_FAST)
        return batch ? pick_spin_v<VAR_FAST, true>(integ, pulse) : pick_spin_v<VAR_FAST, false>(integ, pulse);
    return batch ? pick_spin_v<VAR_EXACT, true>(integ, pulse) : pick_spin_v<VAR_EXACT, false>(integ, pulse);
}
template <int VAR, int INTEG, bool BATCH, bool PULSE>
static const void* spin_sym() {
    if constexpr (use_o4<VAR, INTEG, BATCH, PULSE>())
        return (const void*)&k_spin_o4<VAR, INTEG, BATCH, PULSE>;
    else if constexpr (use_o3<VAR, INTEG, BATCH, PULSE>())
        return (const void*)&k_spin_o3<VAR, INTEG, BATCH, PULSE>;
    else
        return (const void*)&k_spin<VAR, INTEG, BATCH, PULSE>;
}
template <int VAR, bool BATCH>
static const void* spin_symbol_v(int integ, bool pulse) {
    switch (integ) {
    case INT_RK45: return spin_sym<VAR, INT_RK45, BATCH, true>();
    case INT_MAGNUS: return spin_sym<VAR, INT_MAGNUS, BATCH, true>();
    case INT_RK45Q: return spin_sym<VAR, INT_RK45Q, BATCH, true>();
    case INT_RKF45: return spin_sym<VAR, INT_RKF45, BATCH, true>();
    case INT_RKF45Q: return spin_sym<VAR, INT_RKF45Q, BATCH, true>();
    default:
        if constexpr (VAR == VAR_FAST)
            return pulse ? spin_sym<VAR, INT_DOP853, BATCH, true>() : spin_sym<VAR, INT_DOP853, BATCH, false>();
        else
            return spin_sym<VAR, INT_DOP853, BATCH, true>();
    }
}
static const void* spin_symbol(int variant, int integ, bool batch, bool pulse) {
    if (variant == ST_VARIANT_FAST)
        return batch ? spin_symbol_v<VAR_FAST, true>(integ, pulse) : spin_symbol_v<VAR_FAST, false>(integ, pulse);
    return batch ? spin_symbol_v<VAR_EXACT, true>(integ, pulse) : spin_symbol_v<VAR_EXACT, false>(integ, pulse);
}

static int persistent_grid(st_handle* h, const void* sym, int* cache) {
    if (*cache == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        HIPCHK(h, hipGetDevice(&dev));
        HIPCHK(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sym, 256, 0));
        *cache = cus * std::max(per_cu, 1);
    }
    return ST_OK;
}

static RunArgs base_args(st_handle* h) {
    RunArgs R;
    memset(&R, 0, sizeof R);
    R.n = h->n;
    R.work = h->d_work;
    R.queue = h->d_small;
    R.stop_at = h->have_stop_at ? h->d_stop_at : nullptr;
    R.block_bytes = (uint64_t)108 * h->n;
    R.max_events = h->ext.max_events ? h->ext.max_events : ST_DEFAULT_MAX_EVENTS;
    if (h->trace_pid >= 0 && h->d_trace) {
        R.trace = h->d_trace;
        R.trace_count = h->d_trace_count;
        R.trace_pid = (uint32_t)h->trace_pid;
        R.trace_cap = h->trace_cap;
    }
    return R;
}

// ---- pipeline --------------------------------------------------------------
static int ensure_pipeline(st_handle* h) {
    if (h->d_recs) return ST_OK;
    // default budget: 160 GiB, at most half the device memory free now.
    // Every chunk boundary costs a launch drain (lanes idle while the last
    // particles of the chunk finish), which matters when a lane holds only a
    // few particles: 2 x 262,144 particles over 2 s take 3,038 ms in two
    // chunks (2,048 records) and 2,704 ms in one (4,096).
    uint64_t budget = (uint64_t)160 << 30;
    if (h->ext.seg_buffer_mb > 0) {
        budget = (uint64_t)h->ext.seg_buffer_mb << 20;
    } else {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
            budget = std::min<uint64_t>(budget, (uint64_t)(0.5 * (double)free_b));
    }
    const uint64_t n = h->n;
    const uint64_t C_full = budget / (n * sizeof(SegRec));
    // a budget-derived depth stops at 4096 records (seconds of motion per
    // chunk for the densest workloads): small ensembles need no more
    uint64_t C = std::min<uint64_t>(C_full, 4096);
    // Overflow pages for the particles near their end when they reach the
    // depth (k_kin's rule): up to a quarter of the depth's records, on top
    // of the budget by at most a quarter of it.  An explicit chunk depth
    // gets no pool.
    uint64_t pool = 0;
    if (h->ext.chunk_records > 0) C = (uint64_t)h->ext.chunk_records;
    C = std::min<uint64_t>(std::max<uint64_t>(C, 16), 1u << 20);
    if (h->ext.chunk_records == 0) pool = std::min<uint64_t>(C * n / 4, budget / 4 / sizeof(SegRec));
    // Trajectory-only runs read no records: none are kept (a token buffer
    // stands for them), so such a handle holds no record budget and the
    // kinematics run each call in one chunk.
    h->no_recs = h->opt.integratorType == -1;
    if (h->no_recs) {
        C = 16;
        pool = 0;
    }
    h->rec_cap = (uint32_t)C;
    const uint32_t pages = (uint32_t)std::min<uint64_t>(pool / OVF_PAGE, 0xfffffff0u);
    if (pages > 0) {
        HIPCHK(h, hipMalloc(&h->d_ovf, (size_t)pages * OVF_PAGE * sizeof(SegRec)));
        HIPCHK(h, hipMalloc(&h->d_ovf_link, sizeof(uint32_t) * (size_t)pages));
        HIPCHK(h, hipMalloc(&h->d_ovf_first, sizeof(uint32_t) * (size_t)h->n));
        h->ovf_pages = pages;
    }
    HIPCHK(h, hipMalloc(&h->d_recs, (size_t)C * h->n * sizeof(SegRec)));
    HIPCHK(h, hipMalloc(&h->d_list[0], sizeof(uint32_t) * h->n));
    HIPCHK(h, hipMalloc(&h->d_list[1], sizeof(uint32_t) * h->n));
    HIPCHK(h, hipMalloc(&h->d_seg_count, sizeof(uint32_t) * h->n));
    HIPCHK(h, hipMalloc(&h->d_backup, h->state_bytes));
    HIPCHK(h, hipMalloc(&h->d_fail, sizeof(uint32_t) * 3 * (size_t)h->n));
    return ST_OK;
}

static void free_ckpts(st_handle* h) {
    for (auto& c : h->ckpts) {
        if (c.state) (void)hipFree(c.state);
        if (c.list) (void)hipFree(c.list);
    }
    h->ckpts.clear();
}

// Keeps the kinematic state at this chunk's start (k_kin has just backed it
// up for the chunk's particles) and the chunk's active list.
static int save_ckpt(st_handle* h, const uint32_t* list, uint32_t n_list) {
    st_handle::Ckpt c{nullptr, nullptr, n_list};
    HIPCHK(h, hipMalloc(&c.state, h->state_bytes));
    hipError_t e = hipMalloc(&c.list, sizeof(uint32_t) * n_list);
    if (e != hipSuccess) {
        (void)hipFree(c.state);
        return fail(h, ST_E_HIP, "checkpoint allocation failed: %s", hipGetErrorString(e));
    }
    h->ckpts.push_back(c);
    HIPCHK(h, hipMemcpyAsync(c.state, h->d_backup, h->state_bytes, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(c.list, list, sizeof(uint32_t) * n_list, hipMemcpyDeviceToDevice, h->stream));
    return ST_OK;
}

// The chunk loop as a captured HIP graph of K chunks (K even, so the list
// buffers alternate back to list 0), replayed until the device reports an
// empty next list: one host round trip per K chunks.  Kernel shapes are fixed
// (k_kin over all n lanes, k_spin over its persistent grid, k_fixup over 64
// groups); each reads this chunk's list length and failure count on the
// device.  The graph holds the run's arguments, so it is captured per call.
// The kinematics kernel over n_lanes lanes, kbs lanes per workgroup.
static void launch_kin(st_handle* h, const DevState& D, const PipeArgs& A, uint32_t n_lanes, int kbs) {
    const dim3 grid((n_lanes + kbs - 1) / kbs);
    if (h->batch) hipLaunchKernelGGL(k_kin<true>, grid, dim3(kbs), 0, h->stream, h->P, D, A);
    else hipLaunchKernelGGL(k_kin<false>, grid, dim3(kbs), 0, h->stream, h->P, D, A);
}

static int run_pipeline_graph(st_handle* h, PipeArgs A, SpinFn spin, int kbs) {
    const int K = std::min(64, (h->ext.graph_chunks + 1) / 2 * 2);
    hipLaunchKernelGGL(k_graph_start, dim3(1), dim3(1), 0, h->stream, h->d_small, h->n);
    HIPCHK(h, hipGetLastError());
    A.n_list = 0;
    A.n_list_dev = h->d_small + 4;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int c = 0; c < K && e == hipSuccess; ++c) {
        A.list = h->d_list[c & 1];
        A.next_list = h->d_list[(c & 1) ^ 1];
        hipLaunchKernelGGL(k_chunk_head, dim3(1), dim3(1), 0, h->stream, h->d_small);
        launch_kin(h, h->D, A, h->n, kbs);
        e = hipGetLastError();
        if (e == hipSuccess && spin) e = spin(h, dim3(h->grid_spin), A);
        if (e == hipSuccess) {
            if (h->batch)
                hipLaunchKernelGGL(k_fixup<true>, dim3(64), dim3(64), 0, h->stream, h->P, h->D, A, FIXUP_FROM_DEVICE);
            else
                hipLaunchKernelGGL(k_fixup<false>, dim3(64), dim3(64), 0, h->stream, h->P, h->D, A, FIXUP_FROM_DEVICE);
            e = hipGetLastError();
        }
    }
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(h, ST_E_HIP, "chunk graph capture failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : ec));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "chunk graph instantiation failed: %s", hipGetErrorString(e));
    uint32_t left = 0;
    for (;;) {
        e = hipGraphLaunch(x, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&left, h->d_small + 1, sizeof left, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess || left == 0) break;
    }
    (void)hipGraphExecDestroy(x);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "chunk graph replay failed: %s", hipGetErrorString(e));
    return ST_OK;
}

static int run_pipeline(st_handle* h, RunArgs R) {
    int r = ensure_pipeline(h);
    if (r) return r;
    if (!h->ckpts.empty()) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        free_ckpts(h);
    }
    h->ckpt_run = R;
    const int integ = h->opt.integratorType;
    const bool spectral = integ == INT_SPECTRUM;
    if (spectral) {
        const void* sym = h->ext.variant == ST_VARIANT_FAST ? (const void*)&k_spectrum<SPEC_FB, VAR_FAST>
                                                            : (const void*)&k_spectrum<SPEC_FB, VAR_EXACT>;
        if ((r = persistent_grid(h, sym, &h->grid_spec))) return r;
    } else if (integ != INT_NONE &&
               (r = persistent_grid(h, spin_symbol(h->ext.variant, integ, h->batch, h->any_pulse), &h->grid_spin))) {
        return r;
    }
    const int bs = 256;
    // lanes per workgroup of the kinematics kernel (a launch-shape knob)
    const int kbs = h->ext.block_size > 0 ? std::min(256, (h->ext.block_size + 63) / 64 * 64) : 256;
    hipLaunchKernelGGL(k_begin_run, dim3((h->n + bs - 1) / bs), dim3(bs), 0, h->stream, h->D, h->n, h->d_list[0]);
    HIPCHK(h, hipGetLastError());
    PipeArgs A;
    memset(&A, 0, sizeof A);
    A.R = R;
    A.recs = h->d_recs;
    A.C = h->rec_cap;
    A.seg_count = h->d_seg_count;
    A.backup = h->d_backup;
    A.fail_list = h->d_fail;
    A.fail_cap = h->n;
    A.fail_count = h->d_small + 2;
    A.next_count = h->d_small + 1;
    A.integ = integ;
    A.phys = h->d_phys;
    A.no_recs = h->no_recs ? 1u : 0u;
    uint32_t n_list = h->n;
    int cur = 0;
    const SpinFn spin =
        (integ == INT_NONE || spectral) ? nullptr : pick_spin(h->ext.variant, integ, h->batch, h->any_pulse);
    if (h->ext.graph_chunks > 0 && !h->ckpt_on && !h->ktime_on && !spectral)
        return run_pipeline_graph(h, A, spin, kbs);
    if (!spectral && !h->ckpt_on && h->ovf_pages > 0) {
        // the checkpointed (adjoint) and spectral paths keep the fixed layout
        A.ovf = h->d_ovf;
        A.ovf_first = h->d_ovf_first;
        A.ovf_link = h->d_ovf_link;
        A.ovf_head = h->d_small + 3;
        A.ovf_pages = h->ovf_pages;
    }
    while (n_list > 0) {
        A.list = h->d_list[cur];
        A.n_list = n_list;
        A.next_list = h->d_list[cur ^ 1];
        // [0] spin queue, [1] next count, [2] failures, [3] overflow pages taken,
        // [4 + b] queue of spectral block b
        HIPCHK(h, hipMemsetAsync(h->d_small, 0, 8 * sizeof(uint32_t), h->stream));
        if (h->ktime_on) HIPCHK(h, hipEventRecord(h->kev[0], h->stream));
        launch_kin(h, h->D, A, n_list, kbs);
        HIPCHK(h, hipGetLastError());
        if (h->ktime_on) HIPCHK(h, hipEventRecord(h->kev[1], h->stream));
        if (h->ckpt_on && (r = save_ckpt(h, A.list, n_list))) return r;
        if (spin) {
            // the spin kernel only needs its queue head reset; k_kin left the
            // counts in d_small[1..2] and does not touch d_small[0]
            const int grid = std::min<int>(h->grid_spin, (int)((n_list + 255) / 256));
            if (h->ktime_on) HIPCHK(h, hipEventRecord(h->kev[2], h->stream));
            HIPCHK(h, spin(h, dim3(grid), A));
            if (h->ktime_on) HIPCHK(h, hipEventRecord(h->kev[3], h->stream));
        } else if (spectral) {
            // one launch per register block of filters, each with its own queue
            const int grid = std::min<int>(h->grid_spec, (int)((n_list + 255) / 256));
            if (h->ktime_on) HIPCHK(h, hipEventRecord(h->kev[2], h->stream));
            for (int b = 0; b < SPEC_NB; ++b) {
                PipeArgs Ab = A;
                Ab.R.queue = h->d_small + 4 + b;
                SpecArgs S = h->spec;
                S.block = b;
                if (h->ext.variant == ST_VARIANT_FAST)
                    hipLaunchKernelGGL((k_spectrum<SPEC_FB, VAR_FAST>), dim3(grid), dim3(256), 0, h->stream, h->P,
                                       h->D, Ab, S);
                else
                    hipLaunchKernelGGL((k_spectrum<SPEC_FB, VAR_EXACT>), dim3(grid), dim3(256), 0, h->stream, h->P,
                                       h->D, Ab, S);
                HIPCHK(h, hipGetLastError());
            }
            if (h->ktime_on) HIPCHK(h, hipEventRecord(h->kev[3], h->stream));
        }
        uint32_t counts[4];
        HIPCHK(h, hipMemcpyAsync(counts, h->d_small, sizeof counts, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (h->ktime_on) {
            float ms = 0.f;
            HIPCHK(h, hipEventElapsedTime(&ms, h->kev[0], h->kev[1]));
            h->kms[0] += ms;
            h->klaunch[0]++;
            if (spin || spectral) {
                HIPCHK(h, hipEventElapsedTime(&ms, h->kev[2], h->kev[3]));
                h->kms[1] += ms;
                h->klaunch[1]++;
            }
        }
        const uint32_t n_fail = counts[2];
        if (n_fail > 0) {
            if (h->batch)
                hipLaunchKernelGGL(k_fixup<true>, dim3((n_fail + 63) / 64), dim3(64), 0, h->stream, h->P, h->D, A,
                                   n_fail);
            else
                hipLaunchKernelGGL(k_fixup<false>, dim3((n_fail + 63) / 64), dim3(64), 0, h->stream, h->P, h->D, A,
                                   n_fail);
            HIPCHK(h, hipGetLastError());
        }
        n_list = counts[1];
        cur ^= 1;
    }
    return ST_OK;
}

// ---- fused engine ------------------------------------------------------------
static int run_fused(st_handle* h, RunArgs R, bool persist) {
    if (persist) {
        int r = persistent_grid(h, fused_symbol(h->ext.variant, h->opt.integratorType, h->any_pulse), &h->grid_fused);
        if (r) return r;
        HIPCHK(h, hipMemsetAsync(h->d_small, 0, sizeof(uint32_t), h->stream));
        const int grid = std::min<int>(h->grid_fused, (int)((h->n + 255) / 256));
        hipError_t e = pick_fused<true>(h->ext.variant, h->opt.integratorType, h->any_pulse)(h, dim3(grid), dim3(256), R);
        if (e != hipSuccess) return fail(h, ST_E_HIP, "engine launch failed: %s", hipGetErrorString(e));
    } else {
        const int bs = h->ext.block_size > 0 ? std::min(256, (h->ext.block_size + 63) / 64 * 64) : 256;
        hipError_t e = pick_fused<false>(h->ext.variant, h->opt.integratorType, h->any_pulse)(
            h, dim3((h->n + bs - 1) / bs), dim3(bs), R);
        if (e != hipSuccess) return fail(h, ST_E_HIP, "engine launch failed: %s", hipGetErrorString(e));
    }
    return ST_OK;
}

// ---- spectral path (integratorType 6) -----------------------------------------
static int spec_setup(st_handle* h) {
    SpecArgs& S = h->spec;
    memset(&S, 0, sizeof S);
    HIPCHK(h, hipMalloc(&S.agg, sizeof(double) * AGG_FIELDS * 64 * (((size_t)h->n + 63) / 64)));
    HIPCHK(h, hipMalloc(&S.nsamp, sizeof(uint32_t) * h->n));
    HIPCHK(h, hipMalloc(&h->d_spec, sizeof(double) * SPEC_ENTRIES));
    HIPCHK(h, hipMalloc(&h->d_spec_n, sizeof(unsigned long long)));
    HIPCHK(h, hipMalloc(&h->d_spec_part, sizeof(double) * SPEC_RED_BLOCKS * (SPEC_ENTRIES + 1)));
    const st_options& o = h->opt;
    S.na[0] = o.noiseAmplitudes.x; S.na[1] = o.noiseAmplitudes.y; S.na[2] = o.noiseAmplitudes.z;
    S.nw[0] = o.noiseFrequencies.x; S.nw[1] = o.noiseFrequencies.y; S.nw[2] = o.noiseFrequencies.z;
    S.noise_on = (S.na[0] != 0.0 || S.na[1] != 0.0 || S.na[2] != 0.0) ? 1 : 0;
    return ST_OK;
}

// particle::initParticles -> initializeSpectra (include/particle.h:241-242,
// src/particle.cpp:887-891) and SpectrumAggregator::initialize
// (:520-521, 707-708): the Floquet decomposition of the drive, the filter
// constants of its NW frequencies, and empty aggregators and spectrum.
static int spec_init(st_handle* h) {
    floquet::diagonalize(h->P, h->opt.w, h->opt.nmax, h->fd);
    for (int q = 0; q < NW; ++q) {
        const double angle = h->fd.freq[q] * h->fd.dt;
        h->spec.k2[q] = 2 * cos(angle);
        h->spec.cw[q] = cos(angle);
        h->spec.sw[q] = sin(angle);
        h->spec.sw2[q] = 2 * sin(angle);
    }
    hipLaunchKernelGGL(k_spec_clear, dim3((h->n + 255) / 256), dim3(256), 0, h->stream, h->spec, h->n);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemsetAsync(h->d_spec, 0, sizeof(double) * SPEC_ENTRIES, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_spec_n, 0, sizeof(unsigned long long), h->stream));
    return ST_OK;
}

// One interval of the spectral path: clear the aggregators
// (src/particle.cpp:746-747), run the pipeline to the interval's end, and
// add the interval's covariance spectra to the run's (:824-834).
static int spec_interval(st_handle* h, const RunArgs& R) {
    hipLaunchKernelGGL(k_spec_clear, dim3((h->n + 255) / 256), dim3(256), 0, h->stream, h->spec, h->n);
    HIPCHK(h, hipGetLastError());
    const int r = run_pipeline(h, R);
    if (r) return r;
    if (h->ext.spectrum_sum == 0) {
        hipLaunchKernelGGL(k_spec_partial, dim3(1), dim3(64), 0, h->stream, h->spec, h->n, h->d_spec, h->d_spec_n);
    } else {
        const int blocks = std::min<int>(SPEC_RED_BLOCKS, (int)((h->n + 255) / 256));
        hipLaunchKernelGGL(k_spec_all, dim3(blocks), dim3(256), 0, h->stream, h->spec, h->n, h->d_spec_part);
        HIPCHK(h, hipGetLastError());
        hipLaunchKernelGGL(k_spec_final, dim3(1), dim3(320), 0, h->stream, h->d_spec_part, blocks, h->d_spec,
                           h->d_spec_n);
    }
    HIPCHK(h, hipGetLastError());
    return ST_OK;
}

static void fill_floquet(const floquet::Diag& fd, st_floquet* out) {
    memset(out, 0, sizeof *out);
    const Quat m = fd.f_modes_0, e = fd.f_energies;
    const double mm[4] = {m.w, m.x, m.y, m.z}, ee[4] = {e.w, e.x, e.y, e.z};
    memcpy(out->f_modes_0, mm, sizeof mm);
    memcpy(out->f_energies, ee, sizeof ee);
    memcpy(out->frequencies, fd.freq, sizeof fd.freq);
    out->dt = fd.dt;
    out->n_prop = fd.n_prop;
    for (int i = 0; i < fd.n_prop && i < ST_FLOQUET_NPROP; ++i) {
        out->propagators[i][0] = fd.props[i].w;
        out->propagators[i][1] = fd.props[i].x;
        out->propagators[i][2] = fd.props[i].y;
        out->propagators[i][3] = fd.props[i].z;
    }
}

static floquet::Diag read_floquet(const st_floquet* f) {
    floquet::Diag fd;
    fd.f_modes_0 = Quat{f->f_modes_0[0], f->f_modes_0[1], f->f_modes_0[2], f->f_modes_0[3]};
    fd.f_energies = Quat{f->f_energies[0], f->f_energies[1], f->f_energies[2], f->f_energies[3]};
    memcpy(fd.freq, f->frequencies, sizeof fd.freq);
    fd.dt = f->dt;
    fd.n_prop = f->n_prop;
    for (int i = 0; i < f->n_prop; ++i)
        fd.props.push_back(
            Quat{f->propagators[i][0], f->propagators[i][1], f->propagators[i][2], f->propagators[i][3]});
    return fd;
}

extern "C" {

#define ST_STR2(x) #x
#define ST_STR(x) ST_STR2(x)
#if ST_FIELDS_POLYNOMIAL
#define ST_VERSION_FAST ",fast"
#else
#define ST_VERSION_FAST ""
#endif
#ifdef ST_FIELDS_FILE
#define ST_VERSION_FIELDS ST_STR(ST_FIELDS_FILE)
#else
#define ST_VERSION_FIELDS "default"
#endif
ST_API const char* st_version(void) {
    return "spintrack gfx950 exact" ST_VERSION_FAST " dop853,rk45,magnus,rk45q,rkf45,rkf45q,spectral pipeline,fused "
           "fields=" ST_VERSION_FIELDS;
}

ST_API const char* st_last_error(const st_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

static int check_ext(const st_ext* ext, st_ext& x) {
    memset(&x, 0, sizeof x);
    x.pulse_axis = 2;
    if (ext) x = *ext;
    if (x.pulse_axis < 0 || x.pulse_axis > 2) return fail(nullptr, ST_E_ARG, "pulse_axis must be 0, 1 or 2");
    if (x.variant != ST_VARIANT_EXACT && x.variant != ST_VARIANT_FAST)
        return fail(nullptr, ST_E_ARG, "unknown variant %d", x.variant);
    if (!ST_FIELDS_POLYNOMIAL && x.variant == ST_VARIANT_FAST)
        return fail(nullptr, ST_E_UNSUPPORTED,
                    "the field hooks this library was built with (spintrack_fields.h) are not a linear gradient "
                    "and a cos pulse, so the fast variant's per-segment field polynomial does not hold: use "
                    "ST_VARIANT_EXACT");
    if (x.engine != ST_ENGINE_PIPELINE && x.engine != ST_ENGINE_FUSED)
        return fail(nullptr, ST_E_ARG, "unknown engine %d", x.engine);
    if (x.graph_chunks < 0 || x.graph_chunks > 64)
        return fail(nullptr, ST_E_ARG, "graph_chunks must be in 0..64 (got %d)", x.graph_chunks);
    if (x.graph_chunks > 0 && x.engine != ST_ENGINE_PIPELINE)
        return fail(nullptr, ST_E_UNSUPPORTED, "graph_chunks needs the pipeline engine");
    if (x.spectrum_sum != 0 && x.spectrum_sum != 1)
        return fail(nullptr, ST_E_ARG, "spectrum_sum must be 0 or 1 (got %d)", x.spectrum_sum);
    return ST_OK;
}

static int check_options(const st_options& o) {
    if (o.numParticles <= 0) return fail(nullptr, ST_E_ARG, "numParticles must be > 0 (got %d)", o.numParticles);
    if ((uint32_t)o.numParticles > ST_MAX_PARTICLES)
        return fail(nullptr, ST_E_ARG, "numParticles must be at most %u (got %d)", ST_MAX_PARTICLES, o.numParticles);
    if (o.integratorType > 6 || o.integratorType < -1)
        return fail(nullptr, ST_E_UNSUPPORTED,
                    "integratorType %d is not implemented (0 = DOP853, 1 = RK45, 2 = MagnusCFET, 3 = RK45 "
                    "quaternion, 4 = RKF45, 5 = RKF45 quaternion, 6 = noise spectrum / Floquet, -1 = "
                    "trajectory only)",
                    o.integratorType);
    if (o.integratorType == INT_SPECTRUM && !(o.w > 0.0 && std::isfinite(o.w)))
        return fail(nullptr, ST_E_ARG, "integratorType 6 needs a drive frequency w > 0 (got %g)", o.w);
    // the sampling step h: at most 2^40 samples per particle over the run,
    // which also keeps every step well above the rounding of the sample time
    if (o.integratorType == INT_SPECTRUM &&
        !(o.h > 0.0 && std::isfinite(o.h) && std::max(std::fabs(o.t0), std::fabs(o.tf)) / o.h <= 0x1p40))
        return fail(nullptr, ST_E_ARG, "integratorType 6 samples every h: h must be > 0 and >= max(|t0|, |tf|) / 2^40 "
                    "(got h = %g)", o.h);
    return ST_OK;
}

// the fused engine carries DOP853 and RK45 only
static int check_engine_integ(const st_options& o, const st_ext& x) {
    if (x.engine == ST_ENGINE_FUSED && o.integratorType > 1)
        return fail(nullptr, ST_E_UNSUPPORTED, "integratorType %d runs on the pipeline engine only",
                    o.integratorType);
    return ST_OK;
}

// Allocates the device state of n particles; opt is the handle's (run-wide) options.
static int create_common(const st_options& opt, const st_ext& x, uint32_t n, st_handle** out) {
    st_handle* h = new (std::nothrow) st_handle();
    if (!h) return fail(nullptr, ST_E_ARG, "out of host memory");
    h->opt = opt;
    h->opt.numParticles = (int32_t)n;
    h->ext = x;
    h->n = n;
    h->P = derive_params(opt, x);
    h->any_pulse = h->P.ph.pulse_on != 0;
    h->state_bytes = (size_t)field_stride(h->n) * F_COUNT;
    if (hipGetDevice(&h->device) != hipSuccess) h->device = -1;
    hipError_t e = hipMalloc(&h->state, h->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&h->d_work, N_WORK * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&h->d_small, 256);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fail(nullptr, ST_E_HIP, "device set-up failed: %s", hipGetErrorString(e));
        st_destroy(h);
        return ST_E_HIP;
    }
    h->D.base = (char*)h->state;
    h->D.stride = field_stride(h->n);
    h->stream = h->own_stream;
    e = hipMemsetAsync(h->d_work, 0, N_WORK * sizeof(unsigned long long), h->stream);
    if (e != hipSuccess) {
        fail(nullptr, ST_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
        st_destroy(h);
        return ST_E_HIP;
    }
    *out = h;
    return ST_OK;
}

ST_API int st_create(const st_options* opt, const st_ext* ext, st_handle** out) {
    if (!opt || !out) return fail(nullptr, ST_E_ARG, "st_create: null argument");
    *out = nullptr;
    int r = check_options(*opt);
    if (r) return r;
    st_ext x;
    if ((r = check_ext(ext, x))) return r;
    if ((r = check_engine_integ(*opt, x))) return r;
    if ((r = create_common(*opt, x, (uint32_t)opt->numParticles, out))) return r;
    // the spin kernel reads the physics block from device memory
    st_handle* h = *out;
    hipError_t e = hipMalloc(&h->d_phys, sizeof(Phys));
    if (e == hipSuccess) e = hipMemcpy(h->d_phys, &h->P.ph, sizeof(Phys), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fail(nullptr, ST_E_HIP, "device set-up failed: %s", hipGetErrorString(e));
        st_destroy(h);
        *out = nullptr;
        return ST_E_HIP;
    }
    if (opt->integratorType == INT_SPECTRUM && spec_setup(h) != ST_OK) {
        g_create_error = h->err;
        st_destroy(h);
        *out = nullptr;
        return ST_E_HIP;
    }
    return ST_OK;
}

// The options that every configuration of a batch must share: geometry,
// integrator control, time axis and the switches that select kernel code.
static bool same_run(const st_options& a, const st_options& b, const char** what) {
#define ST_SAME(f)                 \
    if (memcmp(&a.f, &b.f, sizeof a.f) != 0) { \
        *what = #f;                \
        return false;              \
    }
    ST_SAME(L) ST_SAME(t0) ST_SAME(tf) ST_SAME(rtol) ST_SAME(atol) ST_SAME(beta) ST_SAME(safe) ST_SAME(fac1)
    ST_SAME(fac2) ST_SAME(hmax) ST_SAME(hmin) ST_SAME(h) ST_SAME(maxPosStep) ST_SAME(ioutInt) ST_SAME(nmax)
    ST_SAME(integratorType) ST_SAME(gas_coll) ST_SAME(gravity) ST_SAME(fixedStepSize) ST_SAME(keepStepSize)
#undef ST_SAME
    return true;
}

ST_API int st_create_batch(const st_options* cfgs, int32_t n_cfg, const st_ext* ext, st_handle** out) {
    if (!cfgs || !out || n_cfg <= 0) return fail(nullptr, ST_E_ARG, "st_create_batch: null argument or n_cfg <= 0");
    *out = nullptr;
    st_ext x;
    int r = check_ext(ext, x);
    if (r) return r;
    if (x.engine != ST_ENGINE_PIPELINE)
        return fail(nullptr, ST_E_UNSUPPORTED, "batched handles run on the pipeline engine only");
    uint64_t total = 0;
    std::vector<uint32_t> off(n_cfg + 1, 0);
    std::vector<uint32_t> seeds(n_cfg);
    std::vector<Phys> phys(n_cfg);
    if (cfgs[0].integratorType == INT_SPECTRUM)
        return fail(nullptr, ST_E_UNSUPPORTED, "integratorType 6 runs on plain handles (one drive per run)");
    for (int c = 0; c < n_cfg; ++c) {
        if ((r = check_options(cfgs[c]))) return r;
        const char* what = "";
        if (!same_run(cfgs[0], cfgs[c], &what))
            return fail(nullptr, ST_E_ARG, "configuration %d differs from configuration 0 in run-wide option '%s'",
                        c, what);
        off[c] = (uint32_t)total;
        total += (uint64_t)cfgs[c].numParticles;
        seeds[c] = cfgs[c].seed;
        phys[c] = derive_phys(cfgs[c], x);
    }
    if (total > ST_MAX_PARTICLES)
        return fail(nullptr, ST_E_ARG, "too many particles in the batch (%llu, at most %u)", (unsigned long long)total,
                    ST_MAX_PARTICLES);
    off[n_cfg] = (uint32_t)total;
    st_handle* h = nullptr;
    if ((r = create_common(cfgs[0], x, (uint32_t)total, &h))) return r;
    h->batch = true;
    h->any_pulse = false;
    for (const Phys& f : phys) h->any_pulse = h->any_pulse || f.pulse_on != 0;
    h->n_cfg = (uint32_t)n_cfg;
    h->cfg_off = off;
    hipError_t e = hipMalloc(&h->d_phys, sizeof(Phys) * n_cfg);
    if (e == hipSuccess) e = hipMalloc(&h->d_cfg_off, sizeof(uint32_t) * (n_cfg + 1));
    if (e == hipSuccess) e = hipMalloc(&h->d_cfg_seed, sizeof(uint32_t) * n_cfg);
    if (e == hipSuccess) e = hipMemcpy(h->d_phys, phys.data(), sizeof(Phys) * n_cfg, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_cfg_off, off.data(), sizeof(uint32_t) * (n_cfg + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_cfg_seed, seeds.data(), sizeof(uint32_t) * n_cfg, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fail(nullptr, ST_E_HIP, "batch set-up failed: %s", hipGetErrorString(e));
        st_destroy(h);
        return ST_E_HIP;
    }
    *out = h;
    return ST_OK;
}

ST_API int st_batch_layout(const st_handle* h, int32_t* n_cfg, uint32_t* offsets) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (n_cfg) *n_cfg = (int32_t)h->n_cfg;
    if (offsets) {
        if (h->batch)
            memcpy(offsets, h->cfg_off.data(), sizeof(uint32_t) * (h->n_cfg + 1));
        else {
            offsets[0] = 0;
            offsets[1] = h->n;
        }
    }
    return ST_OK;
}

ST_API void st_destroy(st_handle* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* bufs[] = {h->state, h->d_targets, h->d_stop_at, h->d_work, h->d_small, h->d_trace, h->d_trace_count,
                    h->d_recs, h->d_ovf, h->d_ovf_first, h->d_ovf_link, h->d_list[0], h->d_list[1],
                    h->d_seg_count, h->d_backup, h->d_fail,
                    h->d_phys, h->d_cfg_off, h->d_cfg_seed, h->d_adj, h->d_replay, h->d_scratch, h->d_adj_recs,
                    h->spec.agg, h->spec.nsamp, h->d_spec, h->d_spec_n, h->d_spec_part};
    free_ckpts(h);
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : h->kev)
        if (e) (void)hipEventDestroy(e);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
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
    if (stride == 0) return fail(h, ST_E_ARG, "stride must be > 0");
    const int bs = 256;
    const dim3 grid((h->n + bs - 1) / bs);
    if (h->batch)
        hipLaunchKernelGGL(k_init<true>, grid, dim3(bs), 0, h->stream, h->P, h->D, h->n, o.seed, first, stride, o.t0,
                           o.tf, h->d_phys, h->d_cfg_off, h->d_cfg_seed, h->n_cfg);
    else
        hipLaunchKernelGGL(k_init<false>, grid, dim3(bs), 0, h->stream, h->P, h->D, h->n, o.seed, first, stride, o.t0,
                           o.tf, nullptr, nullptr, nullptr, 1u);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemsetAsync(h->d_work, 0, N_WORK * sizeof(unsigned long long), h->stream));
    // the pipeline's record buffer (up to tens of GB) is allocated here, with
    // the ensemble, rather than inside the first advance
    if (h->ext.engine == ST_ENGINE_PIPELINE) {
        const int r = ensure_pipeline(h);
        if (r) return r;
    }
    if (o.integratorType == INT_SPECTRUM) {
        const int r = spec_init(h);
        if (r) return r;
    }
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
    const int r = h->opt.integratorType == INT_SPECTRUM ? spec_interval(h, R)
                  : h->ext.engine == ST_ENGINE_FUSED  ? run_fused(h, R, false)
                                                      : run_pipeline(h, R);
    if (r) return r;
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
    if (h->opt.integratorType == INT_SPECTRUM) {
        // the interval sums need every particle at each output time: one
        // pipeline pass per interval, snapshot block k written at target k
        for (int k = 0; k < n_targets; ++k) {
            RunArgs R = base_args(h);
            R.t_single = targets[k];
            R.t_prev = h->last_target;
            R.n_targets = 1;
            R.blocks = blocks_dev ? (char*)blocks_dev + (size_t)k * R.block_bytes : nullptr;
            const int r = spec_interval(h, R);
            if (r) return r;
            h->last_target = targets[k];
        }
        return ST_OK;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_targets, targets, sizeof(double) * n_targets, hipMemcpyHostToDevice, h->stream));
    RunArgs R = base_args(h);
    R.targets = h->d_targets;
    R.t_prev = h->last_target;
    R.n_targets = n_targets;
    R.blocks = (char*)blocks_dev;
    const int r = h->ext.engine == ST_ENGINE_FUSED ? run_fused(h, R, true) : run_pipeline(h, R);
    if (r) return r;
    h->last_target = targets[n_targets - 1];
    return ST_OK;
}

// ---- st_run_noalloc: st_run with its scratch in a static device array ------
// The staging that st_run allocates (the device copy of the output times) and
// the counters its kernels share (work-queue heads, list lengths, failure
// count; the handle's 256-byte counter block otherwise) live in one
// file-scope device array.  Host code never passes `scratch` itself to a
// kernel or a copy: it takes the array's device address per device with
// hipGetSymbolAddress, once, on the first call on that device.
constexpr size_t NOALLOC_COUNTERS = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_COUNTERS + sizeof(double) * ST_NOALLOC_MAX_TARGETS;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[NOALLOC_MAX_DEVICES] = {};

static int scratch_address(st_handle* h, void** out) {
    int dev = -1;
    HIPCHK(h, hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(h, ST_E_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(h, ST_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_addr[dev] = ptr;
    }
    *out = g_scratch_addr[dev];
    return ST_OK;
}

ST_API int st_run_noalloc(st_handle* h, int32_t n_targets, const double* targets, void* blocks_dev) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (!h->initialized) return fail(h, ST_E_STATE, "st_run_noalloc before st_init");
    if (n_targets <= 0 || !targets) return fail(h, ST_E_ARG, "st_run_noalloc needs at least one target");
    if (n_targets > ST_NOALLOC_MAX_TARGETS)
        return fail(h, ST_E_ARG, "st_run_noalloc takes at most %d output times (got %d); use st_run",
                    ST_NOALLOC_MAX_TARGETS, n_targets);
    if (h->ckpt_on)
        return fail(h, ST_E_UNSUPPORTED, "st_run_noalloc does not checkpoint (checkpoints are allocated per run)");
    if (h->device != -1) {
        int dev = -1;
        HIPCHK(h, hipGetDevice(&dev));
        if (dev != h->device)
            return fail(h, ST_E_STATE, "st_run_noalloc: the handle lives on device %d, device %d is current", h->device,
                        dev);
    }
    void* base = nullptr;
    int r = scratch_address(h, &base);
    if (r) return r;
    // the run uses the static array for its counters and staging; the
    // host-driven chunk loop (a captured graph would be instantiated, which
    // allocates) gives the same results as the graph replay
    uint32_t* const own_small = h->d_small;
    double* const own_targets = h->d_targets;
    const int32_t own_cap = h->targets_cap;
    const int32_t own_graph = h->ext.graph_chunks;
    h->d_small = (uint32_t*)base;
    h->d_targets = (double*)((char*)base + NOALLOC_COUNTERS);
    h->targets_cap = ST_NOALLOC_MAX_TARGETS;
    h->ext.graph_chunks = 0;
    // the counters start at zero for this call (the array keeps the last call's values)
    hipError_t e = hipMemsetAsync(base, 0, NOALLOC_COUNTERS, h->stream);
    r = e == hipSuccess ? st_run(h, n_targets, targets, blocks_dev)
                        : fail(h, ST_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
    // the array is shared by every call on this device: finish before returning
    e = hipStreamSynchronize(h->stream);
    h->d_small = own_small;
    h->d_targets = own_targets;
    h->targets_cap = own_cap;
    h->ext.graph_chunks = own_graph;
    if (!r && e != hipSuccess) r = fail(h, ST_E_HIP, "st_run_noalloc: %s", hipGetErrorString(e));
    return r;
}

ST_API int st_set_checkpointing(st_handle* h, int enable) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (enable && h->ext.engine != ST_ENGINE_PIPELINE)
        return fail(h, ST_E_UNSUPPORTED, "checkpointing needs the pipeline engine");
    if (enable && h->opt.integratorType != INT_DOP853)
        return fail(h, ST_E_UNSUPPORTED, "the adjoint integrates the DOP853 path (integratorType 0); got %d",
                    h->opt.integratorType);
    h->ckpt_on = enable != 0;
    return ST_OK;
}

ST_API int st_adjoint(st_handle* h, const void* blocks_dev, const double* seeds, double* grad) {
    if (!h || !blocks_dev || !seeds || !grad) return fail(h, ST_E_ARG, "null argument");
    if (!ST_FIELDS_POLYNOMIAL)
        return fail(h, ST_E_UNSUPPORTED, "st_adjoint differentiates the linear-gradient, cos-pulse field; this "
                    "library was built with other field hooks (spintrack_fields.h)");
    if (h->ckpts.empty()) return fail(h, ST_E_STATE, "st_adjoint needs a checkpointed pipeline run first");
    const RunArgs& R0 = h->ckpt_run;
    const uint32_t K = (uint32_t)R0.n_targets;
    for (size_t i = 0; i < (size_t)3 * K * h->n_cfg; ++i)
        if (!isfinite(seeds[i])) return fail(h, ST_E_ARG, "seed %zu is not finite", i);
    const int bs = 256;
    if (!h->d_adj) HIPCHK(h, hipMalloc(&h->d_adj, (size_t)field_stride(h->n) * A_COUNT));
    if (!h->d_replay) HIPCHK(h, hipMalloc(&h->d_replay, h->state_bytes));
    const size_t scratch_bytes = 256 + sizeof(uint32_t) * (size_t)h->n * 4 + sizeof(double) * 3 * K * h->n_cfg;
    if (h->d_scratch) HIPCHK(h, hipFree(h->d_scratch));
    h->d_scratch = nullptr;
    HIPCHK(h, hipMalloc(&h->d_scratch, scratch_bytes));
    char* sp = h->d_scratch;
    unsigned long long* work = (unsigned long long*)sp;    // 8 counters
    uint32_t* small = (uint32_t*)(sp + 64);                // queue, next_count, fail_count
    uint32_t* next_list = (uint32_t*)(sp + 256);
    uint32_t* fail_list = next_list + h->n;                // 3 per particle
    double* d_seeds = (double*)(fail_list + 3 * (size_t)h->n);
    HIPCHK(h, hipMemcpyAsync(d_seeds, seeds, sizeof(double) * 3 * K * h->n_cfg, hipMemcpyHostToDevice, h->stream));
    const DevState AD{h->d_adj, field_stride(h->n)};
    const DevState DR{h->d_replay, h->D.stride};
    hipLaunchKernelGGL(k_adj_begin, dim3((h->n + bs - 1) / bs), dim3(bs), 0, h->stream, AD, h->n, (int32_t)K - 1,
                       h->P.h0);
    HIPCHK(h, hipGetLastError());
    if (h->batch)   // the replay reads each particle's configuration from its state
        HIPCHK(h, hipMemcpyAsync(DR.f<char>(F_CFG), h->D.f<char>(F_CFG), sizeof(uint64_t) * h->n,
                                 hipMemcpyDeviceToDevice, h->stream));
    PipeArgs A;
    memset(&A, 0, sizeof A);
    A.R = R0;
    A.R.blocks = nullptr;     // the replay rewrites no snapshot block
    A.R.work = work;
    A.R.queue = small;
    A.R.trace = nullptr;
    A.recs = h->d_recs;
    A.C = h->rec_cap;
    A.seg_count = h->d_seg_count;
    A.backup = h->d_backup;
    A.fail_list = fail_list;
    A.fail_cap = h->n;
    A.fail_count = small + 2;
    A.next_count = small + 1;
    A.next_list = next_list;
    A.integ = h->opt.integratorType;
    A.phys = h->d_phys;
    AdjArgs J;
    memset(&J, 0, sizeof J);
    J.recs = h->d_recs;
    J.C = h->rec_cap;
    J.seg_count = h->d_seg_count;
    J.adj = AD;
    J.blocks = (const char*)blocks_dev;
    J.block_bytes = (uint64_t)108 * h->n;
    J.n = h->n;
    J.seeds = d_seeds;
    J.n_targets = K;
    J.phys = h->d_phys;
    // Checkpoint groups: the replay runs G consecutive checkpoint chunks in one
    // stretch, from the first one's start state, into a window of G*C records
    // per particle (the same records the forward run wrote chunk by chunk:
    // its kinematics continue across chunk boundaries unchanged), and one
    // adjoint launch walks the window backwards.  Fewer, longer launches
    // spend less time draining (one particle per lane at this ensemble
    // size, so a launch lasts as long as its slowest particle).  G is what a
    // record window of at most 64 GiB, and a quarter of the free device
    // memory, holds.
    const size_t n_ck = h->ckpts.size();
    uint32_t G = 1;
    {
        size_t free_b = 0, total_b = 0;
        uint64_t budget = (uint64_t)64 << 30;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
            budget = std::min<uint64_t>(budget, (uint64_t)(0.25 * (double)free_b) + h->adj_recs_bytes);
        const uint64_t per_chunk = (uint64_t)h->n * h->rec_cap * sizeof(SegRec);
        G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n_ck, budget / per_chunk));
        if (h->ext.adjoint_window > 0) G = (uint32_t)std::min<uint64_t>(n_ck, h->ext.adjoint_window);
    }
    char* recs = h->d_recs;
    uint32_t* seg_count = h->d_seg_count;
    if (G > 1) {
        const size_t need = (size_t)h->n * G * h->rec_cap * sizeof(SegRec);
        if (h->adj_recs_bytes < need) {
            if (h->d_adj_recs) HIPCHK(h, hipFree(h->d_adj_recs));
            h->d_adj_recs = nullptr;
            h->adj_recs_bytes = 0;
            HIPCHK(h, hipMalloc(&h->d_adj_recs, need));
            h->adj_recs_bytes = need;
        }
        recs = h->d_adj_recs;
    }
    A.recs = recs;
    A.seg_count = seg_count;
    J.recs = recs;
    J.seg_count = seg_count;
    for (size_t hi = n_ck; hi > 0;) {
        const size_t lo = hi >= G ? hi - G : 0;
        const st_handle::Ckpt& c = h->ckpts[lo];
        // replay chunks lo .. hi-1 from chunk lo's start state: the records
        // are bit-identical to the forward run's
        for (int f : KIN_FIELDS)
            HIPCHK(h, hipMemcpyAsync(DR.f<char>(f), c.state + (size_t)f * h->D.stride, h->D.stride,
                                     hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemsetAsync(small, 0, 4 * sizeof(uint32_t), h->stream));
        A.list = c.list;
        A.n_list = c.n_list;
        A.C = (uint32_t)(hi - lo) * h->rec_cap;
        const dim3 grid((c.n_list + bs - 1) / bs);
        launch_kin(h, DR, A, c.n_list, bs);
        HIPCHK(h, hipGetLastError());
        J.list = c.list;
        J.n_list = c.n_list;
        J.C = A.C;   // the window's records per particle: bound and layout stride of both kernels
        J.ck = DevState{c.state, h->D.stride};
        if (h->batch && h->any_pulse)
            hipLaunchKernelGGL((k_adj<true, true>), grid, dim3(bs), 0, h->stream, h->P, h->D, J);
        else if (h->batch)
            hipLaunchKernelGGL((k_adj<true, false>), grid, dim3(bs), 0, h->stream, h->P, h->D, J);
        else if (h->any_pulse)
            hipLaunchKernelGGL((k_adj<false, true>), grid, dim3(bs), 0, h->stream, h->P, h->D, J);
        else
            hipLaunchKernelGGL((k_adj<false, false>), grid, dim3(bs), 0, h->stream, h->P, h->D, J);
        HIPCHK(h, hipGetLastError());
        hi = lo;
    }
    // deterministic reduction: per-block partials per configuration, summed
    // on the host in block order
    uint32_t max_cfg = 0;
    for (uint32_t c = 0; c < h->n_cfg; ++c) {
        const uint32_t lo = h->batch ? h->cfg_off[c] : 0, hi = h->batch ? h->cfg_off[c + 1] : h->n;
        max_cfg = std::max(max_cfg, hi - lo);
    }
    const uint32_t bpc = std::max(1u, std::min<uint32_t>(1024, (max_cfg + 255) / 256));
    double* d_part = nullptr;
    uint32_t* d_off = nullptr;
    std::vector<uint32_t> off = h->batch ? h->cfg_off : std::vector<uint32_t>{0, h->n};
    HIPCHK(h, hipMalloc(&d_part, sizeof(double) * N_GRAD * bpc * h->n_cfg));
    hipError_t e = hipMalloc(&d_off, sizeof(uint32_t) * off.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_adj_partials, dim3(bpc, h->n_cfg), dim3(256), 0, h->stream, AD, d_off, bpc, d_part);
        e = hipGetLastError();
    }
    std::vector<double> part((size_t)N_GRAD * bpc * h->n_cfg);
    std::vector<int64_t> fails(h->n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(fails.data(), AD.f<char>(A_FAIL), sizeof(int64_t) * h->n, hipMemcpyDeviceToHost,
                           h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d_part);
    (void)hipFree(d_off);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "adjoint reduction failed: %s", hipGetErrorString(e));
    for (uint32_t i = 0; i < h->n; ++i)
        if (fails[i]) return fail(h, ST_E_STATE, "adjoint integration failed for particle %u (status %lld)", i,
                                  (long long)fails[i]);
    // gradient = gamma/N * (int u_i r_j, int u x v / c^2, int u)
    std::vector<Phys> phys(h->n_cfg);
    if (h->batch)
        HIPCHK(h, hipMemcpy(phys.data(), h->d_phys, sizeof(Phys) * h->n_cfg, hipMemcpyDeviceToHost));
    else
        phys[0] = h->P.ph;
    for (uint32_t c = 0; c < h->n_cfg; ++c) {
        double s[N_GRAD] = {0};
        for (uint32_t b = 0; b < bpc; ++b)
            for (int f = 0; f < N_GRAD; ++f) s[f] += part[((size_t)c * bpc + b) * N_GRAD + f];
        const double nc = (double)(off[c + 1] - off[c]);
        const double gm = phys[c].gamma / nc;
        for (int f = 0; f < N_GRAD; ++f) grad[c * N_GRAD + f] = gm * s[f] * ((f >= 9 && f < 12) ? IC2 : 1.0);
    }
    return ST_OK;
}

ST_API int st_block_sums(st_handle* h, const void* blocks_dev, int32_t n_blocks, double* out) {
    if (!h || !blocks_dev || !out || n_blocks <= 0) return fail(h, ST_E_ARG, "null argument or n_blocks <= 0");
    std::vector<uint32_t> off = h->batch ? h->cfg_off : std::vector<uint32_t>{0, h->n};
    uint32_t max_cfg = 0;
    for (uint32_t c = 0; c < h->n_cfg; ++c) max_cfg = std::max(max_cfg, off[c + 1] - off[c]);
    const uint32_t wpc = std::max(1u, std::min<uint32_t>(256, (max_cfg + 255) / 256));
    const size_t n_part = (size_t)n_blocks * h->n_cfg * wpc * N_SUMS;
    double* d_part = nullptr;
    uint32_t* d_off = nullptr;
    HIPCHK(h, hipMalloc(&d_part, sizeof(double) * n_part));
    hipError_t e = hipMalloc(&d_off, sizeof(uint32_t) * off.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_block_sums, dim3(wpc, n_blocks, h->n_cfg), dim3(256), 0, h->stream,
                           (const char*)blocks_dev, (uint64_t)108 * h->n, h->n, d_off, wpc, d_part);
        e = hipGetLastError();
    }
    std::vector<double> part(n_part);
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * n_part, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d_part);
    (void)hipFree(d_off);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "block sums failed: %s", hipGetErrorString(e));
    for (int32_t k = 0; k < n_blocks; ++k)
        for (uint32_t c = 0; c < h->n_cfg; ++c) {
            double s[N_SUMS] = {0, 0, 0, 0};
            for (uint32_t w = 0; w < wpc; ++w)
                for (int f = 0; f < N_SUMS; ++f) s[f] += part[(((size_t)k * h->n_cfg + c) * wpc + w) * N_SUMS + f];
            for (int f = 0; f < N_SUMS; ++f) out[((size_t)k * h->n_cfg + c) * N_SUMS + f] = s[f];
        }
    return ST_OK;
}

ST_API int st_phase_stats(st_handle* h, const void* blocks_dev, int32_t n_blocks, double* out, void* phi_dev) {
    if (!h || !blocks_dev || !out || n_blocks <= 0) return fail(h, ST_E_ARG, "null argument or n_blocks <= 0");
    std::vector<uint32_t> off = h->batch ? h->cfg_off : std::vector<uint32_t>{0, h->n};
    double* phi = (double*)phi_dev;
    uint32_t* d_off = nullptr;
    double* d_out = nullptr;
    bool own_phi = false;
    hipError_t e = hipSuccess;
    if (!phi) {
        e = hipMalloc(&phi, sizeof(double) * (size_t)h->n * n_blocks);
        own_phi = e == hipSuccess;
    }
    if (e == hipSuccess) e = hipMalloc(&d_off, sizeof(uint32_t) * off.size());
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * PH_OUT * (size_t)n_blocks * h->n_cfg);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_phase_prep, dim3((h->n + 255) / 256, n_blocks), dim3(256), 0, h->stream,
                           (const char*)blocks_dev, (uint64_t)108 * h->n, h->n, n_blocks, phi);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_phase_stats, dim3(n_blocks, h->n_cfg), dim3(PH_THREADS), 0, h->stream, phi, h->n,
                           d_off, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d_out, sizeof(double) * PH_OUT * (size_t)n_blocks * h->n_cfg, hipMemcpyDeviceToHost,
                           h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (own_phi) (void)hipFree(phi);
    (void)hipFree(d_off);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(h, ST_E_HIP, "phase statistics failed: %s", hipGetErrorString(e));
    return ST_OK;
}

// ---- spectral path entry points ---------------------------------------------
static int need_spectral(st_handle* h) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (h->opt.integratorType != INT_SPECTRUM)
        return fail(h, ST_E_STATE, "not a spectral run (integratorType %d)", h->opt.integratorType);
    if (!h->initialized) return fail(h, ST_E_STATE, "call st_init first");
    return ST_OK;
}

ST_API int st_spectrum_get(st_handle* h, st_spectrum* out) {
    int r = need_spectral(h);
    if (r) return r;
    if (!out) return fail(h, ST_E_ARG, "null argument");
    memset(out, 0, sizeof *out);
    unsigned long long n = 0;
    HIPCHK(h, hipMemcpyAsync(out->variance, h->d_spec, sizeof(double) * SPEC_ENTRIES, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&n, h->d_spec_n, sizeof n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out->n_samples = n;
    memcpy(out->frequencies, h->fd.freq, sizeof out->frequencies);
    out->dt = h->fd.dt;
    return ST_OK;
}

ST_API int st_spectrum_aggregators(st_handle* h, double* agg, uint32_t* n_samples) {
    int r = need_spectral(h);
    if (r) return r;
    if (agg) {
        std::vector<double> dev((size_t)AGG_FIELDS * 64 * (((size_t)h->n + 63) / 64));
        HIPCHK(h, hipMemcpyAsync(dev.data(), h->spec.agg, sizeof(double) * dev.size(), hipMemcpyDeviceToHost,
                                 h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (uint32_t i = 0; i < h->n; ++i)
            for (int f = 0; f < AGG_FIELDS; ++f) agg[(size_t)i * AGG_FIELDS + f] = dev[agg_index(i, f)];
    }
    if (n_samples) {
        HIPCHK(h, hipMemcpyAsync(n_samples, h->spec.nsamp, sizeof(uint32_t) * h->n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return ST_OK;
}

ST_API int st_floquet_info(st_handle* h, st_floquet* out) {
    int r = need_spectral(h);
    if (r) return r;
    if (!out) return fail(h, ST_E_ARG, "null argument");
    fill_floquet(h->fd, out);
    return ST_OK;
}

ST_API int st_floquet_result(st_handle* h, double b_end[3]) {
    int r = need_spectral(h);
    if (r) return r;
    if (!b_end) return fail(h, ST_E_ARG, "null argument");
    st_spectrum sp;
    if ((r = st_spectrum_get(h, &sp))) return r;
    const st_options& o = h->opt;
    const d3 b = floquet::integrate(h->fd, &sp.variance[0][0][0][0], (double)sp.n_samples, h->fd.dt,
                                    d3{o.yi.x, o.yi.y, o.yi.z}, o.t0, o.tf, o.w);
    b_end[0] = b.x; b_end[1] = b.y; b_end[2] = b.z;
    return ST_OK;
}

ST_API int st_floquet_diagonalize(const st_options* opt, const st_ext* ext, st_floquet* out) {
    if (!opt || !out) return fail(nullptr, ST_E_ARG, "st_floquet_diagonalize: null argument");
    st_ext x;
    int r = check_ext(ext, x);
    if (r) return r;
    if (!(opt->w > 0.0 && std::isfinite(opt->w))) return fail(nullptr, ST_E_ARG, "drive frequency w must be > 0");
    floquet::Diag fd;
    floquet::diagonalize(derive_params(*opt, x), opt->w, opt->nmax, fd);
    fill_floquet(fd, out);
    return ST_OK;
}

ST_API int st_floquet_rates(const double f_modes_0[4], const double f_energies[4], const double c_op[4],
                            const double* propagators, int32_t n_prop, double period, const double* frequencies,
                            const double* power, double* Delta, double* X, double* Gamma, double* Zeta,
                            double* Omicron) {
    if (!f_modes_0 || !f_energies || !c_op || !propagators || n_prop <= 0 || !frequencies || !power || !Zeta ||
        !Omicron)
        return fail(nullptr, ST_E_ARG, "st_floquet_rates: null argument or n_prop <= 0");
    using floquet::cd;
    std::vector<Quat> props(n_prop);
    for (int i = 0; i < n_prop; ++i)
        props[i] = Quat{propagators[4 * i], propagators[4 * i + 1], propagators[4 * i + 2], propagators[4 * i + 3]};
    cd pw[NW];
    for (int j = 0; j < NW; ++j) pw[j] = cd(power[2 * j], power[2 * j + 1]);
    double D[2][2][NK];
    cd Xc[2][2][NK], G[2][2][NK], Z[2][2], O[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            Z[i][j] = cd(Zeta[4 * i + 2 * j], Zeta[4 * i + 2 * j + 1]);
            O[i][j] = cd(Omicron[4 * i + 2 * j], Omicron[4 * i + 2 * j + 1]);
        }
    floquet::rates(Quat{f_modes_0[0], f_modes_0[1], f_modes_0[2], f_modes_0[3]},
                   Quat{f_energies[0], f_energies[1], f_energies[2], f_energies[3]},
                   Quat{c_op[0], c_op[1], c_op[2], c_op[3]}, props.data(), n_prop, period, frequencies, pw, D, Xc, G,
                   Z, O);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            Zeta[4 * i + 2 * j] = Z[i][j].real(); Zeta[4 * i + 2 * j + 1] = Z[i][j].imag();
            Omicron[4 * i + 2 * j] = O[i][j].real(); Omicron[4 * i + 2 * j + 1] = O[i][j].imag();
            for (int k = 0; k < NK; ++k) {
                const int e = (2 * i + j) * NK + k;
                if (Delta) Delta[e] = D[i][j][k];
                if (X) { X[2 * e] = Xc[i][j][k].real(); X[2 * e + 1] = Xc[i][j][k].imag(); }
                if (Gamma) { Gamma[2 * e] = G[i][j][k].real(); Gamma[2 * e + 1] = G[i][j][k].imag(); }
            }
        }
    return ST_OK;
}

ST_API int st_floquet_markov(double t0, double tf, double* rho, const double* Zeta, const double* Omicron) {
    if (!rho || !Zeta || !Omicron) return fail(nullptr, ST_E_ARG, "st_floquet_markov: null argument");
    using floquet::cd;
    floquet::M2 m;
    cd Z[2][2], O[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            m.a[i][j] = cd(rho[4 * i + 2 * j], rho[4 * i + 2 * j + 1]);
            Z[i][j] = cd(Zeta[4 * i + 2 * j], Zeta[4 * i + 2 * j + 1]);
            O[i][j] = cd(Omicron[4 * i + 2 * j], Omicron[4 * i + 2 * j + 1]);
        }
    floquet::markov(t0, tf, m, Z, O);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            rho[4 * i + 2 * j] = m.a[i][j].real();
            rho[4 * i + 2 * j + 1] = m.a[i][j].imag();
        }
    return ST_OK;
}

ST_API int st_floquet_integrate(const st_floquet* fd, const st_spectrum* spec, const st_options* opt,
                                double b_end[3]) {
    if (!fd || !spec || !opt || !b_end) return fail(nullptr, ST_E_ARG, "st_floquet_integrate: null argument");
    if (fd->n_prop <= 0 || fd->n_prop > ST_FLOQUET_NPROP) return fail(nullptr, ST_E_ARG, "bad n_prop %d", fd->n_prop);
    const floquet::Diag d = read_floquet(fd);
    const d3 b = floquet::integrate(d, &spec->variance[0][0][0][0], (double)spec->n_samples, spec->dt,
                                    d3{opt->yi.x, opt->yi.y, opt->yi.z}, opt->t0, opt->tf, opt->w);
    b_end[0] = b.x; b_end[1] = b.y; b_end[2] = b.z;
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
    host.resize(h->state_bytes);
    HIPCHK(h, hipMemcpyAsync(host.data(), h->state, h->state_bytes, hipMemcpyDeviceToHost, h->stream));
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
        H.f(F_TEMPH)[i] = q.h;
        H.f(F_TSPIN)[i] = q.t;
        H.f<uint64_t>(F_RX)[i] = q.rng.x; H.f<uint64_t>(F_RY)[i] = q.rng.y;
        H.f<uint64_t>(F_RZ)[i] = q.rng.z; H.f<uint64_t>(F_RW)[i] = q.rng.w;
        H.f(F_SPARE)[i] = q.rng.spare;
        H.f<uint64_t>(F_FLAGS)[i] = pack_flags(q.failure, q.stop != 0, q.rng.hasSpare != 0, (char)q.coll_type,
                                               (char)q.wall_hit);
        H.f<uint64_t>(F_NBOUNCE)[i] = q.n_bounce; H.f<uint64_t>(F_NCOLL)[i] = q.n_coll;
        H.f<uint64_t>(F_NSTEPS)[i] = q.n_steps;
        H.f<uint64_t>(F_PARTID)[i] = q.part_id;
    }
    HIPCHK(h, hipMemcpyAsync(h->state, host.data(), h->state_bytes, hipMemcpyHostToDevice, h->stream));
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

ST_API int st_device_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(nullptr, ST_E_ARG, "null pointer");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(nullptr, ST_E_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return ST_OK;
}

ST_API int st_device_free(void* ptr) {
    if (!ptr) return ST_OK;
    hipError_t e = hipFree(ptr);
    if (e != hipSuccess) return fail(nullptr, ST_E_HIP, "hipFree failed: %s", hipGetErrorString(e));
    return ST_OK;
}

ST_API int st_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(nullptr, ST_E_ARG, "null pointer");
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) return fail(nullptr, ST_E_HIP, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return ST_OK;
}

ST_API int st_host_free(void* ptr) {
    if (!ptr) return ST_OK;
    hipError_t e = hipHostFree(ptr);
    if (e != hipSuccess) return fail(nullptr, ST_E_HIP, "hipHostFree failed: %s", hipGetErrorString(e));
    return ST_OK;
}

ST_API int st_memcpy_d2h(st_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || !dst || !src) return fail(h, ST_E_ARG, "null argument");
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ST_OK;
}

ST_API int st_memcpy_d2d(st_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || !dst || !src) return fail(h, ST_E_ARG, "null argument");
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
    return ST_OK;
}

ST_API int st_timer_start(st_handle* h) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (!h->ev[0]) {
        HIPCHK(h, hipEventCreate(&h->ev[0]));
        HIPCHK(h, hipEventCreate(&h->ev[1]));
    }
    HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    return ST_OK;
}

ST_API int st_timer_stop(st_handle* h, float* ms) {
    if (!h || !ms) return fail(h, ST_E_ARG, "null argument");
    if (!h->ev[0]) return fail(h, ST_E_STATE, "st_timer_stop without st_timer_start");
    HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev[1]));
    HIPCHK(h, hipEventElapsedTime(ms, h->ev[0], h->ev[1]));
    return ST_OK;
}

ST_API int st_kernel_timing(st_handle* h, int enable) {
    if (!h) return fail(nullptr, ST_E_ARG, "null handle");
    if (enable && !h->kev[0])
        for (auto& e : h->kev) HIPCHK(h, hipEventCreate(&e));
    h->ktime_on = enable != 0;
    h->kms[0] = h->kms[1] = 0.0;
    h->klaunch[0] = h->klaunch[1] = 0;
    return ST_OK;
}

ST_API int st_kernel_times(st_handle* h, double ms[2], uint64_t launches[2]) {
    if (!h || !ms || !launches) return fail(h, ST_E_ARG, "null argument");
    for (int j = 0; j < 2; ++j) {
        ms[j] = h->kms[j];
        launches[j] = h->klaunch[j];
    }
    return ST_OK;
}

ST_API int st_work_counters(st_handle* h, uint64_t out[6]) {
    if (!h || !out) return fail(h, ST_E_ARG, "null argument");
    unsigned long long w[N_WORK];
    HIPCHK(h, hipMemcpyAsync(w, h->d_work, sizeof w, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int j = 0; j < 6; ++j) out[j] = w[j];
    return ST_OK;
}

}  // extern "C"
