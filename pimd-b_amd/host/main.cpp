// main.cpp -- pimdb_gpu, the command-line driver (the reference's pimdb,
// src/pimdb.cpp:6-67) running the simulation on one MI355X.
//
//   pimdb_gpu [-in config.ini] [--dim] [--bosonic_alg] [--steps N] [--device-noise] [--factorial]
//
// --factorial runs the factorial bosonic exchange, which the reference
// selects at build time (FACTORIAL_BOSONIC_ALGORITHM).
//
// Errors are printed with the reference's labels; unlike the reference
// (which always exits 0) a failed run exits with status 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "params.hpp"
#include "simulation.hpp"

using namespace pimdb_host;

int main(int argc, char** argv) {
    std::string config = "config.ini";
    RunOptions opts;
    bool info = false;
    try {
        for (int i = 1; i < argc; ++i) {
            if (!std::strcmp(argv[i], "--dim")) {
                std::printf("Program runs 1-, 2- and 3-dimensional systems (the [simulation] key ndim, default 3)\n");
                info = true;
            } else if (!std::strcmp(argv[i], "--bosonic_alg")) {
                std::printf("Program runs the quadratic bosonic algorithm (the factorial one with --factorial).\n");
                info = true;
            } else if (!std::strcmp(argv[i], "-in")) {
                if (i + 1 >= argc) throw std::invalid_argument("-in option requires a filename argument");
                config = argv[++i];
            } else if (!std::strcmp(argv[i], "--steps")) {
                if (i + 1 >= argc) throw std::invalid_argument("--steps option requires a number");
                opts.steps_override = std::atol(argv[++i]);
            } else if (!std::strcmp(argv[i], "--device-noise")) {
                opts.device_noise = true;
            } else if (!std::strcmp(argv[i], "--factorial")) {
                opts.factorial = true;
            }
        }
        if (info) return 0;
        std::printf("[*] Initializing the simulation parameters\n");
        Params params(config);
        Simulation sim(params, opts);
        sim.run();
    } catch (const std::invalid_argument& ex) {
        std::printf("[X] Invalid argument error: %s\n", ex.what());
        return 1;
    } catch (const std::overflow_error& ex) {
        std::printf("[X] Overflow error: %s\n", ex.what());
        return 1;
    } catch (const std::exception& ex) {
        std::printf("[X] Error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
