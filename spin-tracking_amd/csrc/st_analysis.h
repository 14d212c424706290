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

}  // namespace st
