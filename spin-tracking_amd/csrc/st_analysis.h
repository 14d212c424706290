// st_analysis.h -- ensemble reductions over snapshot blocks in HBM.
//
// One workgroup grid per (block, configuration): blockIdx.y = snapshot
// block, blockIdx.z = configuration, blockIdx.x strides over that
// configuration's particles.  Partial sums are written per workgroup and
// added on the host in a fixed order, so results do not depend on
// scheduling.
#pragma once

#include "st_state.h"

namespace st {

constexpr int N_SUMS = 4;   // valid count, sum Sx, sum Sy, sum Sz

// A particle counts when its failure code is 0 and its spin is finite (the
// selection of benchmarks/CheckingFrequencyShiftFromEField.ipynb:29-35).
__global__ __launch_bounds__(256) void k_block_sums(const char* blocks, uint64_t block_bytes, uint32_t n,
                                                    const uint32_t* cfg_off, uint32_t wg_per_cfg, double* partial) {
    const uint32_t k = blockIdx.y, cfg = blockIdx.z;
    const char* B = blocks + (uint64_t)k * block_bytes;
    const double* S = reinterpret_cast<const double*>(B + 56ull * n);
    const int32_t* fail = reinterpret_cast<const int32_t*>(B + 80ull * n);
    const uint32_t lo = cfg_off[cfg], hi = cfg_off[cfg + 1];
    double acc[N_SUMS] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = lo + blockIdx.x * 256 + threadIdx.x; i < hi; i += wg_per_cfg * 256) {
        const double sx = S[3ull * i], sy = S[3ull * i + 1], sz = S[3ull * i + 2];
        if (fail[i] == 0 && __builtin_isfinite(sx) && __builtin_isfinite(sy) && __builtin_isfinite(sz)) {
            acc[0] += 1.0;
            acc[1] += sx;
            acc[2] += sy;
            acc[3] += sz;
        }
    }
    __shared__ double red[256];
    for (int f = 0; f < N_SUMS; ++f) {
        red[threadIdx.x] = acc[f];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[(((uint64_t)k * gridDim.z + cfg) * wg_per_cfg + blockIdx.x) * N_SUMS + f] = red[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Per-particle precession phase statistics, the reference's analysis
// (benchmarks/CheckingFrequencyShiftFromEField.ipynb:5-27 with
// scripts/nedmSim.py:151-178) as device reductions:
//   * particles with failure == 0 and a finite phase in the LAST block are
//     kept (for every block);
//   * phi = atan2(Sz, Sy);
//   * per block: a 1000-bin histogram over [min, max] (numpy's uniform-bin
//     rule and linspace edges), the left edge of its first fullest bin is the
//     reference angle; phases <= ref - pi gain 2 pi, phases >= ref + pi lose
//     2 pi; phases outside mean -+ 5 std (population std) are dropped;
//   * mean (wrapped into [-pi, pi]) and std of the rest.
// ---------------------------------------------------------------------------
constexpr int PH_BINS = 1000;
constexpr int PH_OUT = 6;   // mean (wrapped), std, kept, selected, reference angle, mean before wrapping
constexpr int PH_THREADS = 1024;

__global__ __launch_bounds__(256) void k_phase_prep(const char* blocks, uint64_t block_bytes, uint32_t n,
                                                    int32_t n_blocks, double* phi) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = blockIdx.y;
    if (i >= n) return;
    const char* last = blocks + (uint64_t)(n_blocks - 1) * block_bytes;
    const double* Sl = reinterpret_cast<const double*>(last + 56ull * n) + 3ull * i;
    const int32_t fl = reinterpret_cast<const int32_t*>(last + 80ull * n)[i];
    const bool keep = fl == 0 && !__builtin_isnan(atan2(Sl[2], Sl[1]));
    const double* S = reinterpret_cast<const double*>(blocks + (uint64_t)k * block_bytes + 56ull * n) + 3ull * i;
    phi[(uint64_t)k * n + i] = keep ? atan2(S[2], S[1]) : __builtin_nan("");
}

// block-wide reductions over PH_THREADS threads
__device__ __forceinline__ double ph_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = PH_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}
__device__ __forceinline__ double ph_min(double v, double* red, bool is_max) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = PH_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            red[threadIdx.x] = is_max ? (o > red[threadIdx.x] ? o : red[threadIdx.x])
                                      : (o < red[threadIdx.x] ? o : red[threadIdx.x]);
        }
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// linspace(first, last, PH_BINS + 1)[e]
__device__ __forceinline__ double ph_edge(int e, double first, double last, double step) {
    return e == PH_BINS ? last : (double)e * step + first;
}

__device__ __forceinline__ double ph_shift(double x, double ref) {
    if (x <= ref - PI_D) return x + 2.0 * PI_D;
    if (x >= ref + PI_D) return x - 2.0 * PI_D;
    return x;
}

// one workgroup per (block, configuration)
__global__ __launch_bounds__(PH_THREADS) void k_phase_stats(const double* phi, uint32_t n, const uint32_t* cfg_off,
                                                             double* out) {
    const uint32_t k = blockIdx.x, cfg = blockIdx.y;
    const double* ph = phi + (uint64_t)k * n;
    const uint32_t lo = cfg_off[cfg], hi = cfg_off[cfg + 1];
    __shared__ double red[PH_THREADS];
    __shared__ uint32_t hist[PH_BINS];
    double* o = out + ((uint64_t)k * gridDim.y + cfg) * PH_OUT;
    // range and count of the selected phases
    double mn = __builtin_inf(), mx = -__builtin_inf(), cnt = 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += PH_THREADS) {
        const double x = ph[i];
        if (__builtin_isnan(x)) continue;
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
        cnt += 1.0;
    }
    mn = ph_min(mn, red, false);
    mx = ph_min(mx, red, true);
    cnt = ph_sum(cnt, red);
    if (cnt == 0.0) {
        if (threadIdx.x == 0) {
            const double nan = __builtin_nan("");
            o[0] = nan; o[1] = nan; o[2] = 0.0; o[3] = 0.0; o[4] = nan; o[5] = nan;
        }
        return;
    }
    double first = mn, last = mx;
    if (first == last) {
        first = first - 0.5;
        last = last + 0.5;
    }
    const double denom = last - first;
    const double step = denom / (double)PH_BINS;
    for (int b = threadIdx.x; b < PH_BINS; b += PH_THREADS) hist[b] = 0;
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += PH_THREADS) {
        const double x = ph[i];
        if (__builtin_isnan(x)) continue;
        int idx = (int)(((x - first) / denom) * (double)PH_BINS);
        if (idx == PH_BINS) idx -= 1;
        if (x < ph_edge(idx, first, last, step)) idx -= 1;
        if (x >= ph_edge(idx + 1, first, last, step) && idx != PH_BINS - 1) idx += 1;
        atomicAdd(&hist[idx], 1u);
    }
    __syncthreads();
    // first fullest bin: encode (count, PH_BINS - bin) so that the maximum
    // breaks ties towards the lowest bin
    double best = -1.0;
    for (int b = threadIdx.x; b < PH_BINS; b += PH_THREADS) {
        const double key = (double)hist[b] * 4096.0 + (double)(PH_BINS - b);
        best = key > best ? key : best;
    }
    best = ph_min(best, red, true);
    const int peak = PH_BINS - (int)fmod(best, 4096.0);
    const double ref = ph_edge(peak, first, last, step);
    // aligned phases: mean and population std
    double s = 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += PH_THREADS) {
        const double x = ph[i];
        if (!__builtin_isnan(x)) s += ph_shift(x, ref);
    }
    const double mean = ph_sum(s, red) / cnt;
    s = 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += PH_THREADS) {
        const double x = ph[i];
        if (!__builtin_isnan(x)) {
            const double d = ph_shift(x, ref) - mean;
            s += d * d;
        }
    }
    const double sd = sqrt(ph_sum(s, red) / cnt);
    const double cut_lo = mean - sd * 5.0, cut_hi = mean + sd * 5.0;
    // the phases inside mean -+ 5 std
    double s2 = 0.0, kept = 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += PH_THREADS) {
        const double x = ph[i];
        if (__builtin_isnan(x)) continue;
        const double y = ph_shift(x, ref);
        if (y < cut_lo || y > cut_hi) continue;
        s2 += y;
        kept += 1.0;
    }
    kept = ph_sum(kept, red);
    const double mean2 = ph_sum(s2, red) / kept;
    s2 = 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += PH_THREADS) {
        const double x = ph[i];
        if (__builtin_isnan(x)) continue;
        const double y = ph_shift(x, ref);
        if (y < cut_lo || y > cut_hi) continue;
        s2 += (y - mean2) * (y - mean2);
    }
    const double sd2 = sqrt(ph_sum(s2, red) / kept);
    if (threadIdx.x == 0) {
        double mw = mean2;
        if (mw > PI_D) mw -= PI_D * 2.0;
        if (mw < -PI_D) mw += PI_D * 2.0;
        o[0] = mw; o[1] = sd2; o[2] = kept; o[3] = cnt; o[4] = ref; o[5] = mean2;
    }
}

}  // namespace st
