// run.cpp -- BinaryLogger (src/logger.cpp:8-38) and mainAnalysis
// (src/simulation.cpp:8-62) over the C ABI.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "st_host.h"

namespace sthost {

BinaryLogger::BinaryLogger(const st_options& opt, const char* path) : opt_(opt) { f_ = std::fopen(path, "wb"); }

BinaryLogger::~BinaryLogger() {
    if (f_) std::fclose(f_);
}

void BinaryLogger::writeOptions() { std::fwrite(&opt_, sizeof opt_, 1, f_); }

void BinaryLogger::writeSnapshot(const double* t, const st_coords* pos, const st_coords* v, const st_coords* S,
                                 const int32_t* failure, const uint64_t* n_coll, const uint64_t* n_bounce,
                                 const uint64_t* n_steps) {
    const size_t n = (size_t)opt_.numParticles;
    std::fwrite(t, sizeof *t, n, f_);
    std::fwrite(pos, sizeof *pos, n, f_);
    std::fwrite(v, sizeof *v, n, f_);
    std::fwrite(S, sizeof *S, n, f_);
    std::fwrite(failure, sizeof *failure, n, f_);
    std::fwrite(n_coll, sizeof *n_coll, n, f_);
    std::fwrite(n_bounce, sizeof *n_bounce, n, f_);
    std::fwrite(n_steps, sizeof *n_steps, n, f_);
}

void BinaryLogger::writeBlock(const void* block, size_t bytes) { std::fwrite(block, 1, bytes, f_); }

void BinaryLogger::writeSingle(const char*, const st_coords& value) { std::fwrite(&value, sizeof value, 1, f_); }

namespace {

// particle::postProcess (src/particle.cpp:845-850): a spectral run
// (integratorType 6) ends its output with the Floquet result b_end.
int post_process(st_handle* h, const st_options& opt, Logger& log) {
    if (opt.integratorType != 6) return ST_OK;
    double b[3];
    const int rc = st_floquet_result(h, b);
    if (rc) return rc;
    log.writeSingle("b_end", st_coords{b[0], b[1], b[2]});
    return ST_OK;
}

}  // namespace

namespace {

struct Handle {
    st_handle* h = nullptr;
    ~Handle() { st_destroy(h); }
};

struct DeviceBuf {
    void* p = nullptr;
    ~DeviceBuf() {
        if (p) st_device_free(p);
    }
};

struct HostBuf {
    void* p = nullptr;
    ~HostBuf() {
        if (p) st_host_free(p);
    }
};

// Writes snapshot blocks to the logger on its own thread, one block in
// flight: the device computes interval i+1 while block i goes to the file.
class BlockWriter {
public:
    BlockWriter(Logger& log, size_t bytes) : log_(log), bytes_(bytes), th_([this] { loop(); }) {}
    ~BlockWriter() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    // hands over a block; waits until the previous one is written
    void submit(const void* block) {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [this] { return pending_ == nullptr; });
        pending_ = block;
        cv_.notify_all();
    }
    // waits until every submitted block is written
    void drain() {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [this] { return pending_ == nullptr; });
    }

private:
    void loop() {
        std::unique_lock<std::mutex> g(m_);
        for (;;) {
            cv_.wait(g, [this] { return pending_ != nullptr || stop_; });
            if (pending_) {
                const void* b = pending_;
                g.unlock();
                log_.writeBlock(b, bytes_);
                g.lock();
                pending_ = nullptr;
                cv_.notify_all();
            } else if (stop_) {
                return;
            }
        }
    }
    Logger& log_;
    size_t bytes_;
    std::mutex m_;
    std::condition_variable cv_;
    const void* pending_ = nullptr;
    bool stop_ = false;
    std::thread th_;
};

int report(st_handle* h, int rc, const char* what) {
    if (rc != ST_OK) std::cerr << what << " failed (" << rc << "): " << st_last_error(h) << "\n";
    return rc;
}

// particle::outputData (src/particle.cpp:852-862): copy the eight arrays out
// and hand them to the logger
int output_data(st_handle* h, const st_options& opt, Logger& log) {
    const size_t n = (size_t)opt.numParticles;
    std::vector<double> t(n);
    std::vector<st_coords> pos(n), v(n), S(n);
    std::vector<int32_t> fail(n);
    std::vector<uint64_t> nc(n), nb(n), ns(n);
    const int rc = st_snapshot(h, t.data(), pos.data(), v.data(), S.data(), fail.data(), nc.data(), nb.data(),
                               ns.data());
    if (rc) return report(h, rc, "st_snapshot");
    log.writeSnapshot(t.data(), pos.data(), v.data(), S.data(), fail.data(), nc.data(), nb.data(), ns.data());
    return ST_OK;
}

}  // namespace

int main_analysis(st_options opt, const char* output, const RunConfig& cfg, std::ostream& out) {
    derive(opt);
    if (cfg.tc_override >= 0.0) opt.tc = cfg.tc_override;
    Handle H;
    int rc = st_create(&opt, &cfg.ext, &H.h);
    if (rc) return report(nullptr, rc, "st_create");
    if ((rc = st_init(H.h))) return report(H.h, rc, "st_init");
    std::unique_ptr<Logger> logp;
    try {
        logp = createLogger(opt, output);
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return ST_E_ARG;
    }
    Logger& log = *logp;
    if (!log.ok()) {
        std::cerr << "cannot open " << output << "\n";
        return ST_E_ARG;
    }
    log.writeOptions();
    if ((rc = output_data(H.h, opt, log))) return rc;   // the initial state
    const std::vector<double> times = output_times(opt);
    using clk = std::chrono::high_resolution_clock;
    if (cfg.persistent) {
        // every interval in one call; the blocks come back already laid out
        const size_t bb = st_snapshot_block_bytes(H.h);
        DeviceBuf blocks;
        if ((rc = st_device_alloc(bb * times.size(), &blocks.p))) return report(nullptr, rc, "st_device_alloc");
        const auto t0 = clk::now();
        if ((rc = st_run(H.h, (int32_t)times.size(), times.data(), blocks.p))) return report(H.h, rc, "st_run");
        std::vector<char> host(bb * times.size());
        if ((rc = st_memcpy_d2h(H.h, host.data(), blocks.p, host.size()))) return report(H.h, rc, "st_memcpy_d2h");
        const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t0).count();
        for (size_t k = 0; k < times.size(); ++k) log.writeBlock(host.data() + k * bb, bb);
        if ((rc = post_process(H.h, opt, log))) return report(H.h, rc, "st_floquet_result");
        if (!log.ok()) return report(nullptr, ST_E_ARG, "writing the output file");
        out << "run " << times.size() << " intervals, duration " << (times.empty() ? 0.0 : times.back()) << ", "
            << ms << std::endl;
        return ST_OK;
    }
    // the reference's loop (src/simulation.cpp:84-92): advance, then log the
    // snapshot; each block is built on the device, copied into one of two
    // page-locked buffers and written by the writer thread while the next
    // interval runs.  The file's bytes are those of the synchronous loop.
    const size_t bb = st_snapshot_block_bytes(H.h);
    DeviceBuf dblock;
    HostBuf hb[2];
    if ((rc = st_device_alloc(bb, &dblock.p))) return report(nullptr, rc, "st_device_alloc");
    for (auto& b : hb)
        if ((rc = st_host_alloc(bb, &b.p))) return report(nullptr, rc, "st_host_alloc");
    {
        BlockWriter writer(log, bb);
        for (size_t i = 0; i < times.size(); ++i) {
            const auto t0 = clk::now();
            if ((rc = st_advance_to(H.h, times[i]))) return report(H.h, rc, "st_advance_to");
            if ((rc = st_snapshot_block_device(H.h, dblock.p))) return report(H.h, rc, "st_snapshot_block_device");
            // the writer may still hold the other buffer, never this one
            if ((rc = st_memcpy_d2h(H.h, hb[i & 1].p, dblock.p, bb))) return report(H.h, rc, "st_memcpy_d2h");
            writer.submit(hb[i & 1].p);
            const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t0).count();
            out << "iter " << i << ", duration " << times[i] << ", " << ms << std::endl;
        }
        writer.drain();
    }
    if ((rc = post_process(H.h, opt, log))) return report(H.h, rc, "st_floquet_result");
    if (!log.ok()) return report(nullptr, ST_E_ARG, "writing the output file");
    return ST_OK;
}

}  // namespace sthost
