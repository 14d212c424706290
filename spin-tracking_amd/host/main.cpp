// main.cpp -- command-line driver with the reference's interface
// (main.cpp: `main paramsFile outputFile`, prints the total milliseconds),
// running the simulation on the gfx950 library.
//
//   spintrack_main params.txt out.bin [--persistent] [--fast] [--pulse-axis N]
//                  [--tc SECONDS] [--print-options]
//
// --print-options parses the params file, applies the derived constants and
// writes the 440-byte options block to stdout, without touching the GPU.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "st_host.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cout << "invalid input options\nExpects ./main parameterFile outputFile\n" << std::endl;
        return 1;
    }
    sthost::RunConfig cfg;
    cfg.ext.pulse_axis = 2;
    bool print_options = false;
    for (int a = 3; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--persistent"))
            cfg.persistent = true;
        else if (!std::strcmp(argv[a], "--fast"))
            cfg.ext.variant = ST_VARIANT_FAST;
        else if (!std::strcmp(argv[a], "--pulse-axis") && a + 1 < argc)
            cfg.ext.pulse_axis = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--tc") && a + 1 < argc)
            cfg.tc_override = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--print-options"))
            print_options = true;
        else {
            std::cerr << "unknown argument " << argv[a] << "\n";
            return 1;
        }
    }
    st_options opt;
    try {
        // parser messages go to stderr when stdout carries the options block
        opt = sthost::read_params(argv[1], print_options ? std::cerr : std::cout);
    } catch (const std::exception& e) {
        std::cerr << "params file " << argv[1] << ": " << e.what() << "\n";
        return 1;
    }
    if (print_options) {
        sthost::derive(opt);
        if (cfg.tc_override >= 0.0) opt.tc = cfg.tc_override;
        std::cout.write(reinterpret_cast<const char*>(&opt), sizeof opt);
        return 0;
    }
    const auto t0 = std::chrono::high_resolution_clock::now();
    const int rc = sthost::main_analysis(opt, argv[2], cfg, std::cout);
    const auto ms =
        std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::high_resolution_clock::now() - t0).count();
    std::cout << ms << std::endl;
    return rc == 0 ? 0 : 2;
}
