// hdf5_logger.cpp -- HDF5Logger (src/logger.cpp:40-188) and createLogger
// (:190-207) of the reference, over the HDF5 C library.
//
// File layout, as the reference writes it:
//  * every option as a scalar attribute of the root group: coords as a
//    compound {x, y, z} of doubles (B0, E, L, yi, Gx, Gy, Gz, testNoiseAmp,
//    testNoiseFreq), 24 doubles, 4 ints, 4 bools and dist as unsigned char,
//    nmax and seed as unsigned int (:79-108);
//  * one dataset per snapshot field -- "Time", "Position", "Velocity",
//    "Spin", "Failure State", "Num Collisions", "Num Bounces", "Num Steps" --
//    of shape [nsave, numParticles], nsave = (int)((tf - t0)/ioutInt + 1)
//    (:5), each with an int attribute "nt" counting the rows written (:124-150).
//
// libhdf5 is opened at run time (dlopen of the path the build names, or of
// $SPINTRACK_HDF5_LIB) and only its C entry points are used, so the driver
// carries no link-time dependency on the HDF5 installation or its C++ runtime.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "st_host.h"

#ifdef ST_HDF5
#include <dlfcn.h>
#include <hdf5.h>
#endif

namespace sthost {

void Logger::writeBlock(const void* block, size_t bytes) {
    // split a block of the binary file layout into its eight arrays
    const size_t n = (size_t)numParticles();
    if (bytes != 108 * n) throw std::runtime_error("snapshot block size does not match numParticles");
    const char* p = static_cast<const char*>(block);
    const double* t = reinterpret_cast<const double*>(p);
    const st_coords* pos = reinterpret_cast<const st_coords*>(p + 8 * n);
    const st_coords* v = reinterpret_cast<const st_coords*>(p + 32 * n);
    const st_coords* S = reinterpret_cast<const st_coords*>(p + 56 * n);
    const int32_t* fail = reinterpret_cast<const int32_t*>(p + 80 * n);
    const uint64_t* nc = reinterpret_cast<const uint64_t*>(p + 84 * n);
    const uint64_t* nb = reinterpret_cast<const uint64_t*>(p + 92 * n);
    const uint64_t* ns = reinterpret_cast<const uint64_t*>(p + 100 * n);
    writeSnapshot(t, pos, v, S, fail, nc, nb, ns);
}

#ifdef ST_HDF5

namespace {

// the HDF5 C entry points the logger uses, resolved from the shared library
struct H5Api {
    void* lib = nullptr;
    decltype(&H5open) open = nullptr;
    decltype(&H5Fcreate) Fcreate = nullptr;
    decltype(&H5Fclose) Fclose = nullptr;
    decltype(&H5Screate) Screate = nullptr;
    decltype(&H5Screate_simple) Screate_simple = nullptr;
    decltype(&H5Sselect_hyperslab) Sselect_hyperslab = nullptr;
    decltype(&H5Sclose) Sclose = nullptr;
    decltype(&H5Tcreate) Tcreate = nullptr;
    decltype(&H5Tinsert) Tinsert = nullptr;
    decltype(&H5Tclose) Tclose = nullptr;
    decltype(&H5Acreate2) Acreate2 = nullptr;
    decltype(&H5Aopen) Aopen = nullptr;
    decltype(&H5Awrite) Awrite = nullptr;
    decltype(&H5Aread) Aread = nullptr;
    decltype(&H5Aclose) Aclose = nullptr;
    decltype(&H5Dcreate2) Dcreate2 = nullptr;
    decltype(&H5Dopen2) Dopen2 = nullptr;
    decltype(&H5Dwrite) Dwrite = nullptr;
    decltype(&H5Dclose) Dclose = nullptr;
    decltype(&H5Lexists) Lexists = nullptr;
    hid_t t_double = -1, t_int = -1, t_uint = -1, t_uchar = -1, t_ulong = -1;
};

template <typename F>
void bind(void* lib, F& f, const char* name) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    if (!f) throw std::runtime_error(std::string("libhdf5 lacks ") + name);
}

hid_t native(void* lib, const char* name) {
    void* g = dlsym(lib, name);
    if (!g) throw std::runtime_error(std::string("libhdf5 lacks ") + name);
    return *static_cast<hid_t*>(g);
}

const H5Api& h5() {
    static H5Api api;
    if (api.lib) return api;
    const char* path = std::getenv("SPINTRACK_HDF5_LIB");
    if (!path || !*path) path = ST_HDF5_LIB;
    void* lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!lib) throw std::runtime_error(std::string("cannot load libhdf5 (") + path + "): " + dlerror());
    H5Api a;
    a.lib = lib;
    bind(lib, a.open, "H5open");
    bind(lib, a.Fcreate, "H5Fcreate");
    bind(lib, a.Fclose, "H5Fclose");
    bind(lib, a.Screate, "H5Screate");
    bind(lib, a.Screate_simple, "H5Screate_simple");
    bind(lib, a.Sselect_hyperslab, "H5Sselect_hyperslab");
    bind(lib, a.Sclose, "H5Sclose");
    bind(lib, a.Tcreate, "H5Tcreate");
    bind(lib, a.Tinsert, "H5Tinsert");
    bind(lib, a.Tclose, "H5Tclose");
    bind(lib, a.Acreate2, "H5Acreate2");
    bind(lib, a.Aopen, "H5Aopen");
    bind(lib, a.Awrite, "H5Awrite");
    bind(lib, a.Aread, "H5Aread");
    bind(lib, a.Aclose, "H5Aclose");
    bind(lib, a.Dcreate2, "H5Dcreate2");
    bind(lib, a.Dopen2, "H5Dopen2");
    bind(lib, a.Dwrite, "H5Dwrite");
    bind(lib, a.Dclose, "H5Dclose");
    bind(lib, a.Lexists, "H5Lexists");
    if (a.open() < 0) throw std::runtime_error("H5open failed");
    // the predefined native types are library globals, valid after H5open
    a.t_double = native(lib, "H5T_NATIVE_DOUBLE_g");
    a.t_int = native(lib, "H5T_NATIVE_INT_g");
    a.t_uint = native(lib, "H5T_NATIVE_UINT_g");
    a.t_uchar = native(lib, "H5T_NATIVE_UCHAR_g");
    a.t_ulong = native(lib, "H5T_NATIVE_ULONG_g");
    api = a;
    return api;
}

constexpr unsigned kAccTrunc = 0x0002u;   // H5F_ACC_TRUNC

class HDF5Logger : public Logger {
public:
    HDF5Logger(const st_options& opt, const char* path) : opt_(opt), A_(h5()) {
        nsave_ = (int)((opt.tf - opt.t0) / opt.ioutInt + 1);
        file_ = A_.Fcreate(path, kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
        coords_ = A_.Tcreate(H5T_COMPOUND, sizeof(st_coords));
        A_.Tinsert(coords_, "x", 0, A_.t_double);
        A_.Tinsert(coords_, "y", 8, A_.t_double);
        A_.Tinsert(coords_, "z", 16, A_.t_double);
    }
    ~HDF5Logger() override {
        if (coords_ >= 0) A_.Tclose(coords_);
        if (file_ >= 0) A_.Fclose(file_);
    }
    bool ok() const override { return file_ >= 0 && good_; }
    int32_t numParticles() const override { return opt_.numParticles; }

    void writeOptions() override {
        const char* cn[] = {"B0", "E", "L", "yi", "Gx", "Gy", "Gz", "testNoiseAmp", "testNoiseFreq"};
        const st_coords cv[] = {opt_.B0, opt_.E,  opt_.L,  opt_.yi, opt_.Gx,
                                opt_.Gy, opt_.Gz, opt_.noiseAmplitudes, opt_.noiseFrequencies};
        for (int i = 0; i < 9; ++i) attr(cn[i], coords_, &cv[i]);
        const char* dn[] = {"m",    "t0",   "tf",   "rtol", "atol",     "beta",         "uround",     "safe",
                            "fac1", "fac2", "hmax", "hmin", "h",        "T",            "sqrtKT_m",   "tc",
                            "gamma", "V",   "a",    "w",    "swapStepSize", "maxPosStep", "ioutInt", "diffuse"};
        const double dv[] = {opt_.m,     opt_.t0,     opt_.tf,       opt_.rtol,     opt_.atol,    opt_.beta,
                             opt_.uround, opt_.safe,  opt_.fac1,     opt_.fac2,     opt_.hmax,    opt_.hmin,
                             opt_.h,     opt_.T,      opt_.sqrtKT_m, opt_.tc,       opt_.gamma,   opt_.V,
                             opt_.a,     opt_.w,      opt_.swapStepSize, opt_.maxPosStep, opt_.ioutInt, opt_.diffuse};
        for (int i = 0; i < 24; ++i) attr(dn[i], A_.t_double, &dv[i]);
        const char* in[] = {"integratorType", "numParticles", "numPerGPUBlock", "iout"};
        const int iv[] = {opt_.integratorType, opt_.numParticles, opt_.numPerGPUBlock, opt_.iout};
        for (int i = 0; i < 4; ++i) attr(in[i], A_.t_int, &iv[i]);
        const char* bn[] = {"gas_coll", "gravity", "fixedStepSize", "keepStepSize"};
        const unsigned char bv[] = {(unsigned char)opt_.gas_coll, (unsigned char)opt_.gravity,
                                    (unsigned char)opt_.fixedStepSize, (unsigned char)opt_.keepStepSize};
        for (int i = 0; i < 4; ++i) attr(bn[i], A_.t_uchar, &bv[i]);
        const unsigned char dist = (unsigned char)opt_.dist;
        attr("dist", A_.t_uchar, &dist);
        const unsigned int nmax = opt_.nmax, seed = opt_.seed;
        attr("nmax", A_.t_uint, &nmax);
        attr("seed", A_.t_uint, &seed);
    }

    void writeSnapshot(const double* t, const st_coords* pos, const st_coords* v, const st_coords* S,
                       const int32_t* failure, const uint64_t* n_coll, const uint64_t* n_bounce,
                       const uint64_t* n_steps) override {
        field("Time", A_.t_double, t);
        field("Position", coords_, pos);
        field("Velocity", coords_, v);
        field("Spin", coords_, S);
        field("Failure State", A_.t_int, failure);
        field("Num Collisions", A_.t_ulong, n_coll);
        field("Num Bounces", A_.t_ulong, n_bounce);
        field("Num Steps", A_.t_ulong, n_steps);
    }

    // an attribute of the root group, like the options (writeOption<coords>, :68-70)
    void writeSingle(const char* name, const st_coords& value) override { attr(name, coords_, &value); }

private:
    void attr(const char* name, hid_t type, const void* value) {
        const hid_t sp = A_.Screate(H5S_SCALAR);
        const hid_t a = A_.Acreate2(file_, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
        if (a < 0 || A_.Awrite(a, type, value) < 0) good_ = false;
        if (a >= 0) A_.Aclose(a);
        A_.Sclose(sp);
    }

    // one row of a [nsave, numParticles] dataset, created on first use with
    // its row counter "nt" (:117-150)
    void field(const char* name, hid_t type, const void* data) {
        const hsize_t dims[2] = {(hsize_t)nsave_, (hsize_t)opt_.numParticles};
        if (A_.Lexists(file_, name, H5P_DEFAULT) <= 0) {
            const hid_t fs = A_.Screate_simple(2, dims, nullptr);
            const hid_t d = A_.Dcreate2(file_, name, type, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            A_.Sclose(fs);
            if (d < 0) {
                good_ = false;
                return;
            }
            const hid_t sp = A_.Screate(H5S_SCALAR);
            const hid_t a = A_.Acreate2(d, "nt", A_.t_int, sp, H5P_DEFAULT, H5P_DEFAULT);
            const int zero = 0;
            A_.Awrite(a, A_.t_int, &zero);
            A_.Aclose(a);
            A_.Sclose(sp);
            A_.Dclose(d);
        }
        const hid_t d = A_.Dopen2(file_, name, H5P_DEFAULT);
        const hid_t a = A_.Aopen(d, "nt", H5P_DEFAULT);
        int nt = 0;
        A_.Aread(a, A_.t_int, &nt);
        if (nt >= nsave_) {
            std::cerr << "HDF5Logger: more snapshots than the " << nsave_ << " rows of " << name << "\n";
            good_ = false;
        } else {
            const hid_t fs = A_.Screate_simple(2, dims, nullptr);
            const hsize_t start[2] = {(hsize_t)nt, 0}, count[2] = {1, (hsize_t)opt_.numParticles};
            A_.Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
            const hsize_t mdim[1] = {(hsize_t)opt_.numParticles};
            const hid_t ms = A_.Screate_simple(1, mdim, nullptr);
            if (A_.Dwrite(d, type, ms, fs, H5P_DEFAULT, data) < 0) good_ = false;
            A_.Sclose(ms);
            A_.Sclose(fs);
            ++nt;
            A_.Awrite(a, A_.t_int, &nt);
        }
        A_.Aclose(a);
        A_.Dclose(d);
    }

    st_options opt_;
    const H5Api& A_;
    int nsave_ = 0;
    hid_t file_ = -1, coords_ = -1;
    bool good_ = true;
};

}  // namespace

#endif  // ST_HDF5

// createLogger (src/logger.cpp:190-207): ".hdf5" or ".h5" (any case) selects
// HDF5, anything else the binary format.  Throws std::runtime_error when HDF5
// is asked for and unavailable.
std::unique_ptr<Logger> createLogger(const st_options& opt, const char* path) {
    std::string ext;
    if (const char* dot = std::strrchr(path, '.')) {
        const char* slash = std::strrchr(path, '/');
        if (!slash || dot > slash) ext = dot;
    }
    for (auto& c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext == ".hdf5" || ext == ".h5") {
#ifdef ST_HDF5
        return std::unique_ptr<Logger>(new HDF5Logger(opt, path));
#else
        throw std::runtime_error("HDF5 saving is not enabled, but output file has hdf5 extension.");
#endif
    }
    return std::unique_ptr<Logger>(new BinaryLogger(opt, path));
}

}  // namespace sthost
