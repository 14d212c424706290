// Test program for the host driver's loggers (no GPU): parses a params file,
// then writes K synthetic snapshots through createLogger -- to a binary or an
// HDF5 file by the output path's extension -- as separate arrays, or, with
// "block", as blocks in the binary file layout.
//
//   host_logger_main params.txt out.{bin,h5} K [block]
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "st_host.h"

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    st_options opt = sthost::read_params(argv[1], std::cerr);
    sthost::derive(opt);
    const int K = std::atoi(argv[3]);
    const bool block = argc > 4 && !std::strcmp(argv[4], "block");
    std::unique_ptr<sthost::Logger> log;
    try {
        log = sthost::createLogger(opt, argv[2]);
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 3;
    }
    if (!log->ok()) return 4;
    log->writeOptions();
    const size_t n = (size_t)opt.numParticles;
    std::vector<double> t(n);
    std::vector<st_coords> pos(n), v(n), S(n);
    std::vector<int32_t> fail(n);
    std::vector<uint64_t> nc(n), nb(n), ns(n);
    std::vector<char> blk(108 * n);
    for (int k = 0; k < K; ++k) {
        for (size_t i = 0; i < n; ++i) {
            t[i] = 0.05 * k + 1e-3 * (double)i;
            pos[i] = st_coords{0.01 * k - 0.001 * (double)i, 1.0 / (1.0 + (double)i), -0.25 * k};
            v[i] = st_coords{3.0 + k, -2.0 * (double)i, 1e-7 * (double)(i * k)};
            S[i] = st_coords{0.6, 0.8 * (k % 2 ? -1.0 : 1.0), 1e-300 * (double)i};
            fail[i] = (int32_t)(i % 5) - 2;
            nc[i] = 1000u * (uint64_t)k + i;
            nb[i] = (uint64_t)1 << (i % 63);
            ns[i] = ~(uint64_t)i;
        }
        if (block) {
            char* p = blk.data();
            std::memcpy(p, t.data(), 8 * n);
            std::memcpy(p + 8 * n, pos.data(), 24 * n);
            std::memcpy(p + 32 * n, v.data(), 24 * n);
            std::memcpy(p + 56 * n, S.data(), 24 * n);
            std::memcpy(p + 80 * n, fail.data(), 4 * n);
            std::memcpy(p + 84 * n, nc.data(), 8 * n);
            std::memcpy(p + 92 * n, nb.data(), 8 * n);
            std::memcpy(p + 100 * n, ns.data(), 8 * n);
            log->writeBlock(blk.data(), blk.size());
        } else {
            log->writeSnapshot(t.data(), pos.data(), v.data(), S.data(), fail.data(), nc.data(), nb.data(),
                               ns.data());
        }
    }
    return log->ok() ? 0 : 5;
}
