// simulation.cpp -- see simulation.hpp.
#include "simulation.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>

#include "units.hpp"

namespace pimdb_host {

namespace {

constexpr double kEps = 1.0E-7;   // include/common.h:45

std::string sci(double v, int prec) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.*e", prec, v);
    return buf;
}

// Centred in `width` (extra padding on the right), as the reference's "{:^w}".
std::string centred(const std::string& s, size_t width) {
    if (s.size() >= width) return s;
    const size_t pad = width - s.size(), left = pad / 2;
    return std::string(left, ' ') + s + std::string(pad - left, ' ');
}

std::string report_line(const std::string& name, const std::string& value) {
    std::string n = name;
    if (n.size() < 40) n += std::string(40 - n.size(), ' ');
    return n + "\t:\t" + value + "\n";
}

void make_dir(const std::string& d) {
    if (::mkdir(d.c_str(), 0755) != 0) {
        struct stat st;
        if (::stat(d.c_str(), &st) != 0 || !S_ISDIR(st.st_mode))
            throw std::runtime_error("Cannot create the output folder " + d);
    }
}

template <class T>
T as(const Value& v) {
    return std::get<T>(v);
}

}  // namespace

void check(int status, const char* what) {
    if (status == PIMDB_OK) return;
    const std::string msg = std::string(what);
    switch (status) {
    case PIMDB_E_NONFINITE: throw std::overflow_error(msg);
    case PIMDB_E_ARG: throw std::invalid_argument(msg);
    default: throw std::runtime_error(msg);
    }
}

static void check_ctx(int status, const pimdb_ctx* ctx) {
    if (status == PIMDB_OK) return;
    const char* m = pimdb_last_error(ctx);
    check(status, m ? m : "GPU call failed");
}

// ---------------------------------------------------------------- potentials

void Potential::configure(pimdb_params& p, bool external) const {
    if (external) p.ext_kind = PIMDB_POT_FREE;
    else p.int_kind = PIMDB_POT_FREE;
}

void HarmonicPotential::configure(pimdb_params& p, bool external) const {
    // k = m w^2 with the particle mass for the pair too (simulation.cpp:618-621)
    if (external) {
        p.ext_kind = PIMDB_POT_HARMONIC;
        p.ext_omega = omega_;
    } else {
        p.int_kind = PIMDB_POT_HARMONIC;
        p.int_omega = omega_;
    }
    (void)mass_;
}

void DoubleWellPotential::configure(pimdb_params& p, bool external) const {
    if (!external) throw std::invalid_argument("The double-well potential is an external potential only");
    p.ext_kind = PIMDB_POT_DOUBLE_WELL;
    p.ext_strength = strength_;
    p.ext_location = loc_;
    (void)mass_;
}

void CosinePotential::configure(pimdb_params& p, bool external) const {
    if (!external) throw std::invalid_argument("The cosine potential is an external potential only");
    if (wavelength_ != p.size)
        throw std::invalid_argument("The cosine potential's wavelength must be the system size");
    p.ext_kind = PIMDB_POT_COSINE;
    p.ext_amplitude = amplitude_;
    p.ext_phase = phase_;
}

void DipolePotential::configure(pimdb_params& p, bool external) const {
    if (external) throw std::invalid_argument("The dipole potential is an interaction potential only");
    p.int_kind = PIMDB_POT_DIPOLE;
    p.int_strength = strength_;
}

void AzizPotential::configure(pimdb_params& p, bool external) const {
    if (external) throw std::invalid_argument("The Aziz potential is an interaction potential only");
    p.int_kind = PIMDB_POT_AZIZ;
}

std::unique_ptr<Potential> initializePotential(const std::string& name, const VariantMap& o, double mass,
                                               double size) {
    if (name == "harmonic") return std::make_unique<HarmonicPotential>(mass, as<double>(o.at("omega")));
    if (name == "double_well")
        return std::make_unique<DoubleWellPotential>(mass, as<double>(o.at("strength")), as<double>(o.at("location")));
    if (name == "dipole") return std::make_unique<DipolePotential>(as<double>(o.at("strength")));
    if (name == "cosine")
        return std::make_unique<CosinePotential>(as<double>(o.at("amplitude")), size, as<double>(o.at("phase")));
    if (name == "aziz") return std::make_unique<AzizPotential>();
    return std::make_unique<Potential>();
}

// ----------------------------------------------------------------- exchange

void GpuBosonicExchange::prepare() {
    check_ctx(pimdb_exchange_forward(ctx_, &vb_, nullptr, nullptr, nullptr, nullptr), ctx_);
    check_ctx(pimdb_exchange_estimators(ctx_, &prim_, &p_distinct_, &p_longest_), ctx_);
}

void GpuBosonicExchange::exteriorSpringForce(std::vector<double>& f_first, std::vector<double>& f_last) {
    f_first.assign((size_t)natoms_ * ndim_, 0.0);
    f_last.assign((size_t)natoms_ * ndim_, 0.0);
    check_ctx(pimdb_exchange_adjoint(ctx_, f_first.data(), f_last.data()), ctx_);
}

// --------------------------------------------------------------- simulation

Simulation::Simulation(Params& params, const RunOptions& options) : opts_(options) {
    nbeads = params.get<int>(params.sim, "nbeads");
    ndim = params.get<int>(params.sim, "ndim");
    dt = params.get<double>(params.sim, "dt");
    sfreq = params.get<long>(params.sim, "sfreq");
    steps = opts_.steps_override >= 0 ? opts_.steps_override : params.get<long>(params.sim, "steps");
    threshold = params.get<double>(params.sim, "threshold") * steps;
    bosonic = params.get<bool>(params.sim, "bosonic");
    fixcom = params.get<bool>(params.sim, "fixcom");
    pbc = params.get<bool>(params.sim, "pbc");
    nmthermostat = params.get<bool>(params.sim, "nmthermostat");
    gamma = params.get<double>(params.sim, "gamma");
    temperature = params.get<double>(params.sys, "temperature");
    natoms = params.get<int>(params.sys, "natoms");
    size = params.get<double>(params.sys, "size");
    mass = params.get<double>(params.sys, "mass");
    params_seed = params.get<unsigned int>(params.sim, "seed");
    init_pos_type = params.get<std::string>(params.sim, "init_pos_type");
    init_vel_type = params.get<std::string>(params.sim, "init_vel_type");
    propagator_type = params.get<std::string>(params.sim, "propagator_type");
    thermostat_type = params.get<std::string>(params.sim, "thermostat_type");
    if (sfreq < 1) throw std::invalid_argument("sfreq must be at least one");

    // IPI convention (simulation.cpp:37-52)
    beta = 1.0 / (1.0 * temperature);
    thermo_beta = beta / nbeads;
    omega_p = nbeads / (beta * 1.0);
    spring_constant = mass * omega_p * omega_p;
    bosonic = bosonic && nbeads > 1;   // simulation.cpp:690

    if (propagator_type == "normal_modes") propagator_ = std::make_unique<NormalModesPropagator>();
    else propagator_ = std::make_unique<VelocityVerletPropagator>();
    if (thermostat_type == "langevin") thermostat_ = std::make_unique<LangevinThermostat>(nmthermostat, gamma);
    else if (thermostat_type == "none") thermostat_ = std::make_unique<Thermostat>(nmthermostat);
    else {
        // simulation.cpp:670-681
        const int nchains = params.get<int>(params.sim, "nchains");
        const int kind = thermostat_type == "nose_hoover"      ? PIMDB_THERMO_NOSE_HOOVER
                         : thermostat_type == "nose_hoover_np" ? PIMDB_THERMO_NOSE_HOOVER_NP
                                                               : PIMDB_THERMO_NOSE_HOOVER_NP_DIM;
        thermostat_ = std::make_unique<NoseHooverThermostat>(nmthermostat, kind, nchains);
    }

    external_potential_name = params.get<std::string>(params.external_pot, "name");
    interaction_potential_name = params.get<std::string>(params.interaction_pot, "name");
    int_pot_cutoff = interaction_potential_name == "free" ? 0.0 : params.get<double>(params.interaction_pot, "cutoff");
    if (pbc) int_pot_cutoff = std::min(int_pot_cutoff, 0.5 * size);
    ext_potential_ = initializePotential(external_potential_name, params.external_pot, mass, size);
    int_potential_ = initializePotential(interaction_potential_name, params.interaction_pot, mass, size);

    // ---- the GPU context ----
    gp_.natoms = natoms;
    gp_.nbeads = nbeads;
    gp_.ndim = ndim;
    gp_.mass = mass;
    gp_.temperature = temperature;
    gp_.dt = dt;
    gp_.gamma = gamma;
    gp_.bosonic = bosonic;
    gp_.pbc = pbc;
    gp_.size = size;
    gp_.cutoff = interaction_potential_name == "free" ? 0.0 : params.get<double>(params.interaction_pot, "cutoff");
    gp_.fixcom = fixcom;
    gp_.seed = params_seed;
    gp_.exchange_algo = opts_.factorial ? PIMDB_EXCHANGE_FACTORIAL : PIMDB_EXCHANGE_QUADRATIC;
    ext_potential_->configure(gp_, true);
    int_potential_->configure(gp_, false);
    propagator_->configure(gp_);
    thermostat_->configure(gp_);
    if (!opts_.host_only) {
        check_ctx(pimdb_ctx_create(&gp_, &ctx_), nullptr);
        if (bosonic) exchange_ = std::make_unique<GpuBosonicExchange>(ctx_, natoms, ndim);
    }

    // ---- initial conditions, bead by bead as the reference's ranks ----
    const size_t nb = (size_t)natoms * ndim;
    coord.assign(nb * nbeads, 0.0);
    momenta.assign(nb * nbeads, 0.0);
    forces.assign(nb * nbeads, 0.0);
    for (int b = 0; b < nbeads; ++b) {
        std::mt19937 gen(params_seed + b);   // simulation.cpp:58
        initialize_positions(b, gen, params);
        initialize_momenta(b, gen, params);
        mars_.push_back(std::make_unique<RanMars>((int)(params_seed + b)));   // simulation.cpp:59
    }
    // momenta drawn from Maxwell-Boltzmann get their COM removed on every rank
    if (init_vel_type == "random") zeroMomentum();
    if (!opts_.host_only) {
        push_state();
        open_outputs(params);
    }
}

Simulation::~Simulation() {
    if (ctx_) pimdb_ctx_destroy(ctx_);
}

void Simulation::gen_random_positions(int bead, std::mt19937& gen) {
    std::uniform_real_distribution<double> u_dist(-0.5 * size, 0.5 * size);
    double* x = coord.data() + (size_t)bead * natoms * ndim;
    for (int i = 0; i < natoms; ++i)
        for (int a = 0; a < ndim; ++a) x[i * ndim + a] = u_dist(gen);
}

// Particles at the centres of a grid of boxes (simulation.cpp:133-193).
void Simulation::uniform_particle_grid(int bead) {
    const double volume = std::pow(size, ndim);
    const double init_side = std::pow((1.0 * natoms / volume), -1.0 / (1.0 * ndim));
    int total = 1;
    int count[3] = {1, 1, 1};
    double box[3] = {0, 0, 0};
    for (int a = 0; a < ndim; ++a) {
        count[a] = std::max(static_cast<int>(std::ceil((size / init_side) - kEps)), 1);
        box[a] = size / (1.0 * count[a]);
        total *= count[a];
    }
    if (total < natoms) throw std::runtime_error("Number of grid boxes is less than the number of particles");
    double* x = coord.data() + (size_t)bead * natoms * ndim;
    for (int n = 0; n < natoms; ++n) {
        for (int a = 0; a < ndim; ++a) {
            int scale = 1;
            for (int c = a + 1; c < ndim; ++c) scale *= count[c];
            const int idx = (n / scale) % count[a];
            double pos = (idx + 0.5) * box[a] - 0.5 * size;
            if (pbc) pos -= size * std::floor(pos / size + 0.5);   // common.cpp:41-43
            x[n * ndim + a] = pos;
        }
    }
}

// .xyz reader (common.cpp:49-93): atom count, comment line, "symbol x y z"
// in angstrom.
void Simulation::load_trajectories(const std::string& file, int bead) {
    std::ifstream in(file);
    if (!in.is_open()) throw std::runtime_error("Cannot open the xyz file named " + file + ".");
    int n = 0;
    in >> n;
    if (n != natoms)
        throw std::runtime_error("The number of atoms in the xyz file (" + file +
                                 ") does not match the requested number of atoms.");
    in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    std::string line;
    std::getline(in, line);
    double* x = coord.data() + (size_t)bead * natoms * ndim;
    for (int i = 0; i < natoms; ++i) {
        std::string symbol;
        in >> symbol;
        for (int a = 0; a < ndim; ++a) {
            double v = 0.0;
            in >> v;
            x[i * ndim + a] = to_internal("length", "angstrom", v);
        }
    }
}

// LAMMPS-style velocity file (common.cpp:95-136): count, comment, then
// "index type vx vy vz" in angstrom/ps; momenta = m v.
void Simulation::load_momenta(const std::string& file, int bead) {
    std::ifstream in(file);
    if (!in.is_open()) throw std::runtime_error("Cannot open the velocity file named " + file + ".");
    int n = 0;
    in >> n;
    if (n != natoms)
        throw std::runtime_error("The number of atoms in the velocity file (" + file +
                                 ") does not match the requested number of atoms.");
    in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    std::string line;
    std::getline(in, line);
    double* p = momenta.data() + (size_t)bead * natoms * ndim;
    for (int i = 0; i < natoms; ++i) {
        std::string tok;
        in >> tok >> tok;
        for (int a = 0; a < ndim; ++a) {
            double v = 0.0;
            in >> v;
            p[i * ndim + a] = mass * to_internal("velocity", "angstrom/ps", v);
        }
    }
}

void Simulation::initialize_positions(int bead, std::mt19937& gen, const Params& params) {
    if (init_pos_type == "xyz") {
        load_trajectories(params.get<std::string>(params.sim, "init_pos_xyz_filename"), bead);
    } else if (init_pos_type == "xyz_formatted") {
        const int arg = bead + params.get<int>(params.sim, "init_pos_first_index");
        load_trajectories(format_index(params.get<std::string>(params.sim, "init_pos_xyz_filename_format"), arg), bead);
    } else if (init_pos_type == "grid") {
        uniform_particle_grid(bead);
    } else {
        gen_random_positions(bead, gen);
    }
}

void Simulation::initialize_momenta(int bead, std::mt19937& gen, const Params& params) {
    if (init_vel_type == "manual") {
        load_momenta(format_index("init/vel_{:02}.dat", bead + 1), bead);
    } else if (init_vel_type == "manual_formatted") {
        const int arg = bead + params.get<int>(params.sim, "init_vel_first_index");
        load_momenta(format_index(params.get<std::string>(params.sim, "init_vel_manual_filename_format"), arg), bead);
    } else {
        // Maxwell-Boltzmann at thermo_beta, a new distribution per draw
        // (simulation.cpp:200-217)
        double* p = momenta.data() + (size_t)bead * natoms * ndim;
        for (int i = 0; i < natoms; ++i)
            for (int a = 0; a < ndim; ++a) {
                std::normal_distribution<double> normal(0.0, 1 / std::sqrt(thermo_beta * mass));
                p[i * ndim + a] = mass * normal(gen);
            }
    }
}

// Each bead's sum divided by N*P, then summed over beads (the reference's
// per-rank sums and MPI_Allreduce, simulation.cpp:581-603).
void Simulation::zeroMomentum() {
    std::vector<double> cm(ndim, 0.0);
    const size_t nb = (size_t)natoms * ndim;
    for (int b = 0; b < nbeads; ++b) {
        std::vector<double> part(ndim, 0.0);
        const double* p = momenta.data() + nb * b;
        for (int i = 0; i < natoms; ++i)
            for (int a = 0; a < ndim; ++a) part[a] += p[i * ndim + a];
        for (int a = 0; a < ndim; ++a) cm[a] += part[a] / (natoms * nbeads);
    }
    for (int b = 0; b < nbeads; ++b) {
        double* p = momenta.data() + nb * b;
        for (int i = 0; i < natoms; ++i)
            for (int a = 0; a < ndim; ++a) p[i * ndim + a] -= cm[a];
    }
}

void Simulation::push_state() { check_ctx(pimdb_set_state(ctx_, coord.data(), momenta.data()), ctx_); }

void Simulation::pull_state() {
    check_ctx(pimdb_get_state(ctx_, coord.data(), momenta.data(), forces.data()), ctx_);
}

void Simulation::draw_noise(long nsteps, std::vector<double>& out) {
    const size_t nb = (size_t)natoms * ndim;
    out.resize((size_t)nsteps * 2 * nbeads * nb);
    size_t e = 0;
    for (long s = 0; s < nsteps; ++s)
        for (int o = 0; o < 2; ++o)
            for (int b = 0; b < nbeads; ++b)
                for (size_t k = 0; k < nb; ++k) out[e++] = mars_[b]->gaussian();
}

// ----------------------------------------------------------------- outputs

void Simulation::open_outputs(const Params& params) {
    make_dir(opts_.output_dir);
    struct Spec {
        const char* key;
        const char* kind;
        const char* family;
        const char* file;
    };
    const Spec specs[] = {{"positions", "position", "length", "position_%d.xyz"},
                          {"velocities", "velocity", "velocity", "velocity_%d.dat"},
                          {"forces", "force", "force", "force_%d.dat"}};
    for (const Spec& s : specs) {
        const std::string units = params.states.at(s.key);
        if (units == "off" || units == "false") continue;
        StateFile sf;
        sf.kind = s.kind;
        sf.unit = (units == "on" || units == "true" || units == "none") ? "atomic_unit" : units;
        try {
            to_user(s.family, sf.unit, 1.0);
        } catch (const std::invalid_argument&) {
            throw std::invalid_argument(std::string("Invalid output unit for ") + s.kind + " state.");
        }
        for (int b = 0; b < nbeads; ++b) {
            char name[64];
            std::snprintf(name, sizeof(name), s.file, b);
            sf.files.emplace_back(opts_.output_dir + "/" + name, std::ios::out | std::ios::app);
        }
        states_.push_back(std::move(sf));
    }

    // observables (simulation.cpp:801-815) and their labels (energy.cpp:10-19,
    // classical.cpp:11-19, bosonic.cpp:9-12)
    auto unit_of = [&](const std::string& key) {
        const std::string u = params.observables.at(key);
        return u == "none" ? std::string() : u;
    };
    if (params.observables.at("energy") != "off") {
        const std::string u = unit_of("energy");
        quantities_.push_back({"kinetic", "energy", u, 0});
        const bool ext_free = external_potential_name == "free", int_free = interaction_potential_name == "free";
        if (!ext_free || !int_free) {
            quantities_.push_back({"potential", "energy", u, 1});
            if (!ext_free && !int_free) {
                quantities_.push_back({"ext_pot", "energy", u, 2});
                quantities_.push_back({"int_pot", "energy", u, 3});
            }
            quantities_.push_back({"virial", "energy", u, 4});
        }
    }
    if (params.observables.at("classical") != "off") {
        const std::string u = unit_of("classical");
        quantities_.push_back({"temperature", "temperature", "kelvin", 5});
        quantities_.push_back({"cl_kinetic", "energy", u, 6});
        quantities_.push_back({"cl_spring", "energy", u, 7});
        // classical.cpp:13-15, 24-26
        if (thermostat_type.rfind("nose_hoover", 0) == 0) quantities_.push_back({"nh_energy", "energy", u, 10});
    }
    if (bosonic && params.observables.at("bosonic") != "off") {
        quantities_.push_back({"prob_dist", "number", "", 8});
        quantities_.push_back({"prob_all", "number", "", 9});
    }
    if (params.observables.at("gsf") != "off")
        throw std::invalid_argument("The gsf observable is not available on the GPU path");
    for (const Quantity& q : quantities_) to_user(q.family, q.unit, 1.0);   // unknown units fail now

    obs_file_.open(opts_.output_dir + "/simulation.out", std::ios::out | std::ios::app);
    if (!obs_file_.is_open()) throw std::runtime_error("Failed to open simulation.out.");
    obs_file_ << centred("step", 16);
    for (const Quantity& q : quantities_) obs_file_ << " " << centred(q.label, 16);
    obs_file_ << '\n';
}

// Position, velocity and force files, one per bead, at the start of a step
// (src/states/{position,velocity,force}.cpp).
void Simulation::write_states(long step) {
    if (states_.empty()) return;
    pull_state();
    const size_t nb = (size_t)natoms * ndim;
    const char* pad = ndim == 1 ? " 0.0 0.0" : (ndim == 2 ? " 0.0" : "");
    for (StateFile& sf : states_) {
        for (int b = 0; b < nbeads; ++b) {
            std::ostringstream out;
            out << natoms << "\n" << "Step " << step << "\n";
            for (int i = 0; i < natoms; ++i) {
                if (sf.kind == "position") out << "1";
                else out << (i + 1) << " 1";
                for (int a = 0; a < ndim; ++a) {
                    const size_t e = nb * b + (size_t)i * ndim + a;
                    double v;
                    if (sf.kind == "position") v = to_user("length", sf.unit, coord[e]);
                    else if (sf.kind == "velocity") v = to_user("velocity", sf.unit, momenta[e] / mass);
                    else v = to_user("force", sf.unit, forces[e]);
                    out << " " << centred(sci(v, 12), 20);
                }
                out << pad << "\n";
            }
            sf.files[b] << out.str();
        }
    }
}

void Simulation::log_observables(long step) {
    if (quantities_.empty()) {
        obs_file_ << centred(sci((double)step, 8), 16) << '\n';
        return;
    }
    double out[11];
    check_ctx(pimdb_observables(ctx_, out), ctx_);
    obs_file_ << centred(sci((double)step, 8), 16);
    for (const Quantity& q : quantities_) obs_file_ << " " << centred(sci(to_user(q.family, q.unit, out[q.index]), 8), 16);
    obs_file_ << '\n';
}

void Simulation::write_report(double wall_time) const {
    std::ofstream rep(opts_.output_dir + "/report.txt", std::ios::out | std::ios::app);
    rep << "---------\nParameters\n---------\n";
    if (bosonic) {
        rep << report_line("Statistics", "Bosonic");
        rep << report_line("Bosonic algorithm", opts_.factorial ? "Naive" : "Feldman-Hirshberg");
    } else {
        rep << report_line("Statistics", "Boltzmannonic");
    }
    rep << report_line("Time propagation algorithm", propagator_type);
    rep << report_line("Periodic boundary conditions", pbc ? "true" : "false");
    rep << report_line("Dimension", std::to_string(ndim));
    rep << report_line("Seed", std::to_string(params_seed));
    rep << report_line("Coordinate initialization method", init_pos_type);
    rep << report_line("Number of atoms", std::to_string(natoms));
    rep << report_line("Number of beads", std::to_string(nbeads));
    rep << report_line("Total number of MD steps", std::to_string(steps));
    rep << report_line("Interaction potential name", interaction_potential_name);
    rep << report_line("External potential name", external_potential_name);
    rep << report_line("Noise", opts_.device_noise ? "device (Philox4x32-10)" : "host RanMars streams");
    rep << "---------\n";
    rep << report_line("Wall time (sec)", sci(wall_time, 5));
    rep << report_line("Wall time per step (sec)", sci(wall_time / std::max(1L, steps), 5));
}

// OBABO loop of the reference (simulation.cpp:222-290): states at the start
// of a step, the step itself on the GPU, observables of production steps at
// the save frequency.  Runs of steps between outputs are one pimdb_step call.
void Simulation::run() {
    if (!ctx_) throw std::runtime_error("Simulation was created without a GPU context");
    std::printf("[*] Running the simulation\n");
    const auto t0 = std::chrono::steady_clock::now();
    const bool host_noise = !opts_.device_noise && thermostat_->draws_noise();
    if (opts_.device_noise) check_ctx(pimdb_graph_capture(ctx_), ctx_);
    std::vector<double> noise;
    // host noise is uploaded in chunks of at most about 32 MB
    const long per_step = 2L * nbeads * natoms * ndim;
    const long chunk = std::max(1L, (4L << 20) / per_step);
    long step = 0;
    while (step <= steps) {
        if (step % sfreq == 0) write_states(step);
        // run up to (and including) the next step whose end is logged or
        // whose start is written
        long next = std::min(steps, (step / sfreq + 1) * sfreq - 1);
        if (step % sfreq == 0 && step >= threshold) next = step;
        for (long s0 = step; s0 <= next;) {
            const long n = host_noise ? std::min(chunk, next - s0 + 1) : next - s0 + 1;
            if (host_noise) draw_noise(n, noise);
            check_ctx(pimdb_step(ctx_, (int)n, host_noise ? noise.data() : nullptr), ctx_);
            s0 += n;
        }
        if (next % sfreq == 0 && next >= threshold) log_observables(next);
        step = next + 1;
    }
    obs_file_.flush();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("[*] Simulation finished running successfully (Runtime = %.3g sec)\n", wall);
    write_report(wall);
}

}  // namespace pimdb_host
