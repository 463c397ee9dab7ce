// host_capi.cpp -- extern "C" entry points of include/pimdb_host.h.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "params.hpp"
#include "pimdb_host.h"
#include "random_mars.hpp"
#include "simulation.hpp"
#include "units.hpp"

using namespace pimdb_host;

struct pimdb_host_ranmars {
    RanMars gen;
    explicit pimdb_host_ranmars(int seed) : gen(seed) {}
};

namespace {

void put(char* err, int errlen, const std::string& msg) {
    if (!err || errlen <= 0) return;
    std::strncpy(err, msg.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = '\0';
}

// Runs f, mapping the reference's exception types to status codes.
template <class F>
int guarded(char* err, int errlen, F f) {
    try {
        f();
        put(err, errlen, "");
        return PIMDB_OK;
    } catch (const std::invalid_argument& e) {
        put(err, errlen, e.what());
        return PIMDB_E_ARG;
    } catch (const std::overflow_error& e) {
        put(err, errlen, e.what());
        return PIMDB_E_NONFINITE;
    } catch (const std::exception& e) {
        put(err, errlen, e.what());
        return PIMDB_E_HIP;
    }
}

}  // namespace

extern "C" {

int pimdb_host_to_internal(const char* family, const char* unit, double number, double* out, char* err, int errlen) {
    if (!family || !unit || !out) return PIMDB_E_ARG;
    return guarded(err, errlen, [&] { *out = to_internal(family, unit, number); });
}

pimdb_host_ranmars* pimdb_host_ranmars_new(int seed) {
    try {
        return new pimdb_host_ranmars(seed);
    } catch (const std::exception&) {
        return nullptr;
    }
}
double pimdb_host_ranmars_uniform(pimdb_host_ranmars* r) { return r->gen.uniform(); }
double pimdb_host_ranmars_gaussian(pimdb_host_ranmars* r) { return r->gen.gaussian(); }
void pimdb_host_ranmars_free(pimdb_host_ranmars* r) { delete r; }

int pimdb_host_load_config(const char* path, pimdb_params* out, long* steps, long* sfreq, char* err, int errlen) {
    if (!path || !out) return PIMDB_E_ARG;
    return guarded(err, errlen, [&] {
        Params params(path);
        RunOptions o;
        o.host_only = true;
        Simulation sim(params, o);
        *out = sim.gpu_params();
        if (steps) *steps = sim.steps;
        if (sfreq) *sfreq = sim.sfreq;
    });
}

int pimdb_host_initial_state(const char* path, double* x, double* p, char* err, int errlen) {
    if (!path) return PIMDB_E_ARG;
    return guarded(err, errlen, [&] {
        Params params(path);
        RunOptions o;
        o.host_only = true;
        Simulation sim(params, o);
        if (x) std::memcpy(x, sim.coord.data(), sizeof(double) * sim.coord.size());
        if (p) std::memcpy(p, sim.momenta.data(), sizeof(double) * sim.momenta.size());
    });
}

int pimdb_host_draw_noise(const char* path, long nsteps, double* out, char* err, int errlen) {
    if (!path || !out || nsteps < 0) return PIMDB_E_ARG;
    return guarded(err, errlen, [&] {
        Params params(path);
        RunOptions o;
        o.host_only = true;
        Simulation sim(params, o);
        std::vector<double> v;
        sim.draw_noise(nsteps, v);
        std::memcpy(out, v.data(), sizeof(double) * v.size());
    });
}

int pimdb_host_run(const char* path, long steps, int flags, const char* output_dir, char* err, int errlen) {
    if (!path) return PIMDB_E_ARG;
    return guarded(err, errlen, [&] {
        Params params(path);
        RunOptions o;
        o.steps_override = steps;
        o.device_noise = (flags & PIMDB_HOST_DEVICE_NOISE) != 0;
        o.factorial = (flags & PIMDB_HOST_FACTORIAL) != 0;
        if (output_dir) o.output_dir = output_dir;
        Simulation sim(params, o);
        sim.run();
    });
}

}  // extern "C"
