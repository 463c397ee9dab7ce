// simulation.hpp -- the host side of the GPU path: the reference's
// Simulation, BosonicExchange, Potential, Propagator and Thermostat surfaces
// over the C ABI of include/pimdb_gpu.h.
//
// One process owns the whole ring polymer (the reference runs one MPI rank
// per bead).  What stays on the host is what the reference does once or
// between steps: configuration (Params), initial conditions with the
// reference's generators, the per-bead RanMars noise streams, unit
// conversion and the output files.  Every MD step, force and observable is
// computed on the GPU.
#pragma once

#include <fstream>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "params.hpp"
#include "pimdb_gpu.h"
#include "random_mars.hpp"

namespace pimdb_host {

// Throws the exception type the reference throws for the same condition:
// PIMDB_E_NONFINITE -> std::overflow_error (quadratic_bosonic_exchange.cpp:
// 92-97), PIMDB_E_ARG -> std::invalid_argument, anything else ->
// std::runtime_error.
void check(int status, const char* what);

// ---- potentials (include/potentials/*.h; factory simulation.cpp:612-645) ----
// On the GPU path a potential is a description that selects and parameterises
// a kernel: configure() writes it into pimdb_params.
class Potential {
public:
    virtual ~Potential() = default;
    virtual void configure(pimdb_params& p, bool external) const;   // free
};
class HarmonicPotential : public Potential {
public:
    HarmonicPotential(double mass, double omega) : mass_(mass), omega_(omega) {}
    void configure(pimdb_params& p, bool external) const override;
private:
    double mass_, omega_;
};
class DoubleWellPotential : public Potential {
public:
    DoubleWellPotential(double mass, double strength, double loc) : mass_(mass), strength_(strength), loc_(loc) {}
    void configure(pimdb_params& p, bool external) const override;
private:
    double mass_, strength_, loc_;
};
class CosinePotential : public Potential {
public:
    CosinePotential(double amplitude, double wavelength, double phase)
        : amplitude_(amplitude), wavelength_(wavelength), phase_(phase) {}
    void configure(pimdb_params& p, bool external) const override;
private:
    double amplitude_, wavelength_, phase_;
};
class DipolePotential : public Potential {
public:
    explicit DipolePotential(double strength) : strength_(strength) {}
    void configure(pimdb_params& p, bool external) const override;
private:
    double strength_;
};
class AzizPotential : public Potential {
public:
    void configure(pimdb_params& p, bool external) const override;
};
std::unique_ptr<Potential> initializePotential(const std::string& name, const VariantMap& options, double mass,
                                               double size);

// ---- propagators and thermostats (propagator.h, thermostat.h) ----
// The GPU step runs the whole OBABO sequence; these select its variants.
class Propagator {
public:
    virtual ~Propagator() = default;
    virtual void configure(pimdb_params& p) const = 0;
};
class VelocityVerletPropagator : public Propagator {
public:
    void configure(pimdb_params& p) const override { p.propagator = PIMDB_PROP_CARTESIAN; }
};
class NormalModesPropagator : public Propagator {
public:
    void configure(pimdb_params& p) const override { p.propagator = PIMDB_PROP_NORMAL_MODES; }
};
class Thermostat {   // "none" (thermostat.cpp: step() does nothing)
public:
    explicit Thermostat(bool nmthermostat) : nm_(nmthermostat) {}
    virtual ~Thermostat() = default;
    virtual void configure(pimdb_params& p) const { p.thermostat = PIMDB_THERMO_NONE; p.nmthermostat = nm_; }
    virtual bool draws_noise() const { return false; }
protected:
    bool nm_;
};
class LangevinThermostat : public Thermostat {
public:
    LangevinThermostat(bool nmthermostat, double gamma) : Thermostat(nmthermostat), gamma_(gamma) {}
    void configure(pimdb_params& p) const override {
        p.thermostat = PIMDB_THERMO_LANGEVIN;
        p.nmthermostat = nm_;
        p.gamma = gamma_;
    }
    bool draws_noise() const override { return true; }
private:
    double gamma_;
};
// NoseHooverThermostat and its per-particle / per-degree-of-freedom variants
// (nose_hoover.cpp); the chains live in the GPU context.
class NoseHooverThermostat : public Thermostat {
public:
    NoseHooverThermostat(bool nmthermostat, int kind, int nchains) : Thermostat(nmthermostat), kind_(kind), nchains_(nchains) {}
    void configure(pimdb_params& p) const override {
        p.thermostat = kind_;
        p.nmthermostat = nm_;
        p.nchains = nchains_;
    }
private:
    int kind_, nchains_;
};

// ---- the exchange (bosonic_exchange_base.h:16-54) on a GPU context ----
class GpuBosonicExchange {
public:
    explicit GpuBosonicExchange(pimdb_ctx* ctx, int natoms, int ndim) : ctx_(ctx), natoms_(natoms), ndim_(ndim) {}
    void prepare();                       // forward pass from the context's current beads 1 and P
    double effectivePotential() const { return vb_; }
    double primEstimator() const { return prim_; }
    double getDistinctProbability() const { return p_distinct_; }
    double getLongestProbability() const { return p_longest_; }
    // exterior spring forces of beads 1 and P, N x D each ([particle][axis])
    void exteriorSpringForce(std::vector<double>& f_first, std::vector<double>& f_last);

private:
    pimdb_ctx* ctx_;
    int natoms_, ndim_;
    double vb_ = 0.0, prim_ = 0.0, p_distinct_ = 0.0, p_longest_ = 0.0;
};

// Options of a run that the configuration file does not carry.
struct RunOptions {
    bool device_noise = false;   // Philox noise on the GPU instead of the RanMars streams
    bool factorial = false;      // the factorial bosonic exchange (FACTORIAL_BOSONIC_ALGORITHM)
    long steps_override = -1;    // replaces [simulation] steps when >= 0
    std::string output_dir = "output";
    // initial conditions and noise streams only: no GPU context, no files
    bool host_only = false;
};

// ---- Simulation (include/simulation.h, src/simulation.cpp) ----
class Simulation {
public:
    Simulation(Params& params, const RunOptions& options = RunOptions());
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    void run();

    // state of all beads, host layout [bead][particle][axis]
    std::vector<double> coord, momenta, forces;
    void pull_state();   // device -> coord, momenta, forces
    void push_state();   // coord, momenta -> device

    void zeroMomentum();   // on the host arrays (simulation.cpp:581-603)

    // parameters, as the reference's fields
    double temperature = 0, beta = 0, thermo_beta = 0, dt = 0, size = 0, gamma = 0, threshold = 0;
    double mass = 0, spring_constant = 0, omega_p = 0;
    int natoms = 0, nbeads = 0, ndim = 3;
    long sfreq = 1000, steps = 0;
    bool bosonic = false, fixcom = true, pbc = false, nmthermostat = false;
    unsigned int params_seed = 0;
    std::string init_pos_type, init_vel_type, propagator_type, thermostat_type;
    std::string external_potential_name, interaction_potential_name;
    double int_pot_cutoff = 0.0;

    pimdb_ctx* context() const { return ctx_; }
    GpuBosonicExchange* bosonic_exchange() { return exchange_.get(); }
    const pimdb_params& gpu_params() const { return gp_; }

    // Gaussian deviates for `nsteps` steps in the reference's consumption
    // order: [step][O-step][bead or mode][particle][axis], bead b from its
    // own RanMars(seed + b).
    void draw_noise(long nsteps, std::vector<double>& out);

private:
    void initialize_positions(int bead, std::mt19937& gen, const Params& params);
    void initialize_momenta(int bead, std::mt19937& gen, const Params& params);
    void gen_random_positions(int bead, std::mt19937& gen);
    void uniform_particle_grid(int bead);
    void load_trajectories(const std::string& file, int bead);
    void load_momenta(const std::string& file, int bead);

    void write_states(long step);
    void open_outputs(const Params& params);
    void log_observables(long step);
    void write_report(double wall_time) const;

    RunOptions opts_;
    pimdb_params gp_{};
    pimdb_ctx* ctx_ = nullptr;
    std::unique_ptr<GpuBosonicExchange> exchange_;
    std::unique_ptr<Potential> ext_potential_, int_potential_;
    std::unique_ptr<Propagator> propagator_;
    std::unique_ptr<Thermostat> thermostat_;
    std::vector<std::unique_ptr<RanMars>> mars_;

    struct StateFile {
        std::string kind, unit;
        std::vector<std::ofstream> files;   // one per bead
    };
    std::vector<StateFile> states_;
    struct Quantity {
        std::string label, family, unit;
        int index;   // into pimdb_observables' output
    };
    std::vector<Quantity> quantities_;
    std::ofstream obs_file_;
};

}  // namespace pimdb_host
