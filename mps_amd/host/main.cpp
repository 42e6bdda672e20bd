// main.cpp — the reference's command line (src/main.cpp:12-42):
//   mps -s|--setting <settings.yml>
// Pressure-solver options beyond the reference come from the environment:
//   MPS_AMD_PRECISION=fp64|mixed  MPS_AMD_TOLERANCE=<rel residual>
//   MPS_AMD_WARM_START=0|1
#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>

#include "simulation.hpp"

namespace fs = std::filesystem;

int main(int argc, char** argv) {
    std::string setting;
    for (int k = 1; k < argc; k++) {
        const std::string a = argv[k];
        if (a == "-s" || a == "--setting") {
            if (k + 1 >= argc) {
                std::cout << "Too few arguments for '" << a << "'." << std::endl;
                std::exit(EXIT_FAILURE);
            }
            setting = argv[++k];
            if (!fs::exists(setting)) {
                std::cout << "ERROR: Input YAML file cannot be found in specified path: " + setting << std::endl;
                std::exit(EXIT_FAILURE);
            }
            std::cout << "Setting file: " << setting << std::endl;
        } else if (a == "-h" || a == "--help") {
            std::cout << "Usage: mps [--help] --setting VAR\n\nOptional arguments:\n  -h, --help     shows help "
                         "message and exits\n  -s, --setting  path to YAML file for settings. [required]"
                      << std::endl;
            return 0;
        } else {
            std::cout << "Unknown argument: " << a << std::endl;
            std::exit(EXIT_FAILURE);
        }
    }
    if (setting.empty()) {
        std::cout << "-s: required." << std::endl;
        std::exit(EXIT_FAILURE);
    }

    Simulation simulation;
    mps_solver_options o{};
    o.precision = MPS_PREC_FP64;
    o.tolerance = DBL_EPSILON;
    o.inner_tolerance = 1e-4;
    o.check_interval = 32;
    o.use_graph = 1;
    bool custom = false;
    if (const char* p = std::getenv("MPS_AMD_PRECISION")) {
        o.precision = std::strcmp(p, "mixed") == 0 ? MPS_PREC_MIXED : MPS_PREC_FP64;
        custom = true;
    }
    if (const char* t = std::getenv("MPS_AMD_TOLERANCE")) {
        o.tolerance = std::strtod(t, nullptr);
        custom = true;
    }
    if (const char* w = std::getenv("MPS_AMD_WARM_START")) {
        const int v = std::atoi(w);
        o.warm_start = v == 2 ? 2 : (v ? 1 : 0);  // 2: extrapolated from the last two solutions
        custom = true;
    }
    if (const char* m = std::getenv("MPS_AMD_PRECONDITIONER")) {
        o.preconditioner = std::strcmp(m, "multigrid") == 0 ? MPS_PRECOND_MULTIGRID : MPS_PRECOND_JACOBI;
        custom = true;
    }
    if (custom) simulation.setSolver(o);
    try {
        simulation.run(fs::path(setting));
    } catch (const std::exception& e) {
        std::cout << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    return 0;
}
