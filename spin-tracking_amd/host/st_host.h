// st_host.h -- C++ host side of the spin-tracking drop-in: the params.txt
// grammar, the binary snapshot logger and the mainAnalysis run loop, over
// the C ABI of include/spintrack.h.  Mirrors the reference's host code
// (src/optionsParser.cpp, src/logger.cpp, src/simulation.cpp) so an existing
// run script keeps working against the gfx950 library.
#pragma once

#include <cstdio>
#include <iosfwd>
#include <string>
#include <memory>
#include <vector>

#include "../../include/spintrack.h"

namespace sthost {

// include/options.h:11-60 defaults.
st_options default_options();

// One line of a params file (src/optionsParser.cpp:40-175): "key, v1[, v2,
// v3]", whitespace removed; a line with fewer than two elements and an
// unknown key are reported on `log` and ignored.  Malformed numbers throw
// std::invalid_argument, as std::stod / std::stoi do in the reference.
void parse_line(st_options& opt, const std::string& line, std::ostream& log);

// src/optionsParser.cpp:177-197: defaults, then every line of the file; a
// file that cannot be opened leaves the defaults.
st_options read_params(const char* path, std::ostream& log);

// src/simulation.cpp:9-11: tc = 1.6e-4*m/(k*T^8), sqrtKT_m = sqrt(k*T/m).
void derive(st_options& opt);

// floor((tf - t0)/ioutInt) output times (i+1)*ioutInt (src/simulation.cpp:79-85).
std::vector<double> output_times(const st_options& opt);

// Logger (include/logger.h:12-30): the options, then one snapshot per
// output time.
class Logger {
public:
    virtual ~Logger() = default;
    virtual bool ok() const = 0;
    virtual int32_t numParticles() const = 0;
    virtual void writeOptions() = 0;
    virtual void writeSnapshot(const double* t, const st_coords* pos, const st_coords* v, const st_coords* S,
                               const int32_t* failure, const uint64_t* n_coll, const uint64_t* n_bounce,
                               const uint64_t* n_steps) = 0;
    // a block in the binary file layout (st_snapshot_block_bytes bytes:
    // t, pos, v, S, failure, n_coll, n_bounce, n_steps); by default split
    // into its arrays and written as a snapshot
    virtual void writeBlock(const void* block, size_t bytes);
    // one named vector after the snapshots (writeSingle, src/logger.cpp:31-33,
    // 68-70): the spectral run's b_end
    virtual void writeSingle(const char* name, const st_coords& value) = 0;
};

// BinaryLogger (src/logger.cpp:8-38): the 440-byte options block, then one
// snapshot block per output -- t, pos, v, S, failure, n_coll, n_bounce,
// n_steps, numParticles entries each.
class BinaryLogger : public Logger {
public:
    BinaryLogger(const st_options& opt, const char* path);
    ~BinaryLogger() override;
    bool ok() const override { return f_ != nullptr; }
    int32_t numParticles() const override { return opt_.numParticles; }
    void writeOptions() override;
    void writeSnapshot(const double* t, const st_coords* pos, const st_coords* v, const st_coords* S,
                       const int32_t* failure, const uint64_t* n_coll, const uint64_t* n_bounce,
                       const uint64_t* n_steps) override;
    // the block is already the file's bytes
    void writeBlock(const void* block, size_t bytes) override;
    // the raw 24 bytes, no name
    void writeSingle(const char* name, const st_coords& value) override;

private:
    st_options opt_;
    FILE* f_ = nullptr;
};

// createLogger (src/logger.cpp:190-207): HDF5Logger for a ".h5" / ".hdf5"
// path (any case), BinaryLogger otherwise.  Throws std::runtime_error if
// HDF5 output is asked for and this build or host has no libhdf5.
std::unique_ptr<Logger> createLogger(const st_options& opt, const char* path);

struct RunConfig {
    st_ext ext{};
    bool persistent = false;   // one st_run over all output times instead of one call per interval
    double tc_override = -1;   // >= 0: replaces the derived gas mean free time
};

// mainAnalysis (src/simulation.cpp:8-62): derive tc / sqrtKT_m, create and
// seed the ensemble, log the initial snapshot, then advance to every output
// time and log it, printing "iter i, duration t, ms" per interval.  Returns
// 0, or the failing st_* status (its message on stderr).
int main_analysis(st_options opt, const char* output, const RunConfig& cfg, std::ostream& out);

}  // namespace sthost
