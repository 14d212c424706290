// params.cpp -- the reference's params.txt grammar (src/optionsParser.cpp:11-197)
// and the derived constants of src/simulation.cpp:9-11, over st_options.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <ostream>
#include <sstream>

#include "st_host.h"

namespace sthost {

st_options default_options() {
    st_options o{};
    o.B0 = {3e-6, 0.0, 0.0};
    o.E = {75e5, 0.0, 0.0};
    o.L = {0.07, 0.1, 0.4};
    o.yi = {0.0, 0.0, 1.0};
    o.Gx = o.Gy = o.Gz = {0.0, 0.0, 0.0};
    o.m = 1.20239e-26;
    o.t0 = 0.0;
    o.tf = 10.0;
    o.rtol = 1e-12;
    o.atol = 1e-12;
    o.beta = 0.0;
    o.uround = 1e-16;
    o.safe = 0.9;
    o.fac1 = 0.333;
    o.fac2 = 6.0;
    o.hmax = 1.0;
    o.hmin = 1.0e-8;
    o.h = 0.001;
    o.T = 4.2;
    o.sqrtKT_m = 0.0;
    o.tc = 0.0;
    o.gamma = -2.038e8;
    o.V = 5.0;
    o.a = 0.0;
    o.w = 0.0;
    o.swapStepSize = 1.0 - 4;   // sic (include/options.h:39)
    o.maxPosStep = 0.1;
    o.ioutInt = 0.05;
    o.diffuse = 1.0;
    o.nmax = 10000000;
    o.seed = 0;
    o.integratorType = 0;
    o.numParticles = 1000;
    o.numPerGPUBlock = 128;
    o.iout = 2;
    o.dist = 'C';
    o.gas_coll = 1;
    o.gravity = 1;
    o.fixedStepSize = 0;
    o.keepStepSize = 0;
    o.noiseAmplitudes = o.noiseFrequencies = {0.0, 0.0, 0.0};
    return o;
}

namespace {

// comma-separated fields, each with all whitespace removed
std::vector<std::string> fields_of(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : line) {
        if (c == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (!std::isspace((unsigned char)c)) {
            cur.push_back(c);
        }
    }
    out.push_back(cur);
    return out;
}

// std::boolalpha extraction after lower-casing: only "true" reads as true
uint8_t bool_of(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    std::istringstream is(s);
    bool b = false;
    is >> std::boolalpha >> b;
    return b ? 1 : 0;
}

st_coords coords_of(const std::vector<std::string>& e) {
    return st_coords{std::stod(e.at(1)), std::stod(e.at(2)), std::stod(e.at(3))};
}

}  // namespace

void parse_line(st_options& opt, const std::string& line, std::ostream& log) {
    const std::vector<std::string> e = fields_of(line);
    if (e.size() < 2) {
        log << "Misunderstood input line in options file\n\t" << line << "\n";
        return;
    }
    const std::string& k = e[0];
    struct D {
        const char* key;
        double st_options::*field;
    };
    static const D doubles[] = {
        {"m", &st_options::m},         {"T", &st_options::T},       {"diffuse", &st_options::diffuse},
        {"gamma", &st_options::gamma}, {"V", &st_options::V},       {"a", &st_options::a},
        {"w", &st_options::w},         {"t0", &st_options::t0},     {"tf", &st_options::tf},
        {"rtol", &st_options::rtol},   {"atol", &st_options::atol}, {"beta", &st_options::beta},
        {"uround", &st_options::uround}, {"safe", &st_options::safe}, {"fac1", &st_options::fac1},
        {"fac2", &st_options::fac2},   {"hmax", &st_options::hmax}, {"hmin", &st_options::hmin},
        {"h", &st_options::h},         {"maxPosStep", &st_options::maxPosStep},
        {"swapStepSize", &st_options::swapStepSize}, {"ioutInt", &st_options::ioutInt}};
    struct Cd {
        const char* key;
        st_coords st_options::*field;
    };
    static const Cd coords[] = {{"L", &st_options::L},   {"yi", &st_options::yi}, {"Gx", &st_options::Gx},
                                {"Gy", &st_options::Gy}, {"Gz", &st_options::Gz}, {"B0", &st_options::B0},
                                {"E", &st_options::E}};
    struct B {
        const char* key;
        uint8_t st_options::*field;
    };
    static const B bools[] = {{"gas_coll", &st_options::gas_coll},
                              {"gravity", &st_options::gravity},
                              {"fixedStepSize", &st_options::fixedStepSize},
                              {"keepStepSize", &st_options::keepStepSize}};
    for (const Cd& c : coords)
        if (k == c.key) {
            opt.*(c.field) = coords_of(e);
            return;
        }
    for (const D& d : doubles)
        if (k == d.key) {
            opt.*(d.field) = std::stod(e[1]);
            return;
        }
    for (const B& b : bools)
        if (k == b.key) {
            opt.*(b.field) = bool_of(e[1]);
            return;
        }
    if (k == "dist") {
        opt.dist = e[1].empty() ? '\0' : e[1][0];
    } else if (k == "nmax") {
        opt.nmax = (uint32_t)std::stoi(e[1]);
    } else if (k == "seed") {
        opt.seed = (uint32_t)std::stoi(e[1]);
    } else if (k == "integratorType") {
        opt.integratorType = std::stoi(e[1]);
    } else if (k == "numParticles") {
        opt.numParticles = std::stoi(e[1]);
    } else if (k == "numPerGPUBlock") {
        opt.numPerGPUBlock = std::stoi(e[1]);
    } else if (k == "iout") {
        opt.iout = std::stoi(e[1]);
    } else {
        log << "Unrecognized Option: " << k << "\nLine Ignored\n";
    }
}

st_options read_params(const char* path, std::ostream& log) {
    st_options o = default_options();
    std::ifstream f(path);
    std::string line;
    while (f.is_open() && std::getline(f, line)) parse_line(o, line, log);
    return o;
}

void derive(st_options& opt) {
    const double k = 1.380649e-23;
    opt.tc = 1.6e-4 * opt.m / (k * std::pow(opt.T, 8.0));
    opt.sqrtKT_m = std::sqrt(k * opt.T / opt.m);
}

std::vector<double> output_times(const st_options& opt) {
    const int n = (int)std::floor((opt.tf - opt.t0) / opt.ioutInt);
    std::vector<double> t;
    for (int i = 0; i < n; ++i) t.push_back(((double)i + 1.0) * opt.ioutInt);
    return t;
}

}  // namespace sthost
