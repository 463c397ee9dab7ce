// params.cpp -- see params.hpp.
#include "params.hpp"

#include <algorithm>
#include <cstdio>
#include <regex>
#include <stdexcept>
#include <sys/stat.h>
#include <vector>

#include "units.hpp"

namespace pimdb_host {
namespace {

bool in_list(const std::string& s, const std::vector<std::string>& list) {
    return std::find(list.begin(), list.end(), s) != list.end();
}

bool file_exists(const std::string& path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

std::string fmt3(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%4.3f", v);
    return buf;
}

}  // namespace

bool is_index_pattern(const std::string& p) {
    static const std::regex re(R"(\{(:0?\d*d?)?\})");
    return std::regex_search(p, re);
}

std::string format_index(const std::string& pattern, int value) {
    static const std::regex re(R"(\{(:(0?)(\d*)d?)?\})");
    std::string out;
    auto begin = pattern.cbegin();
    std::smatch m;
    while (std::regex_search(begin, pattern.cend(), m, re)) {
        out.append(begin, m[0].first);
        char spec[16] = "%d";
        if (m[1].matched) {
            const bool zero = m[2].length() > 0;
            const int width = m[3].length() ? std::stoi(m[3].str()) : 0;
            std::snprintf(spec, sizeof(spec), zero ? "%%0%dd" : "%%%dd", width);
        }
        char buf[64];
        std::snprintf(buf, sizeof(buf), spec, value);
        out += buf;
        begin = m[0].second;
    }
    out.append(begin, pattern.cend());
    return out;
}

bool Params::parse_token_parentheses(const std::string& input, std::string& token, std::string& value) {
    static const std::regex pattern(R"(\s*(\w+)\((.*)\)\s*|\s*(\w+)\s*)");
    std::smatch m;
    if (std::regex_match(input, m, pattern) && (m[1].matched || m[3].matched)) {
        token = m[1].matched ? m[1].str() : m[3].str();
        value = m[2].str();
        return true;
    }
    return false;
}

Params::Params(const std::string& filename) : reader(filename) {
    if (reader.parse_error() < 0) throw std::invalid_argument("Unable to read the configuration file " + filename);
    load();
}

Params Params::from_string(const std::string& text) {
    Params p;
    p.reader = IniReader::from_string(text);
    p.load();
    return p;
}

void Params::load() {
    using namespace sections;
    const IniReader& r = reader;

    // ---- simulation (params.cpp:14-35) ----
    sim["dt"] = quantity("time", r.get(SIMULATION, "dt", "1.0 femtosecond"));
    sim["threshold"] = r.get_real(SIMULATION, "threshold", 0.1);
    sim["gamma"] = r.get_real(SIMULATION, "gamma", -1.0);
    if (std::get<double>(sim["gamma"]) < 0) sim["gamma"] = 1 / (100.0 * std::get<double>(sim["dt"]));
    sim["nchains"] = r.get_integer(SIMULATION, "nchains", 4);
    if (const int nc = std::get<int>(sim["nchains"]); nc < 1)
        throw std::invalid_argument("The specified number of Nose-Hoover chains (" + std::to_string(nc) +
                                    ") is less than one!");
    sim["steps"] = static_cast<long>(std::stod(r.get(SIMULATION, "steps", "1e5")));
    sim["sfreq"] = r.get_long(SIMULATION, "sfreq", 1000);
    sim["nbeads"] = r.get_integer(SIMULATION, "nbeads", 4);
    if (const int nb = std::get<int>(sim["nbeads"]); nb < 1)
        throw std::invalid_argument("The specified number of beads (" + std::to_string(nb) + ") is less than one!");
    sim["seed"] = static_cast<unsigned int>(std::stod(r.get(SIMULATION, "seed", "1234")));
    // spatial dimension: compile-time in the reference (NDIM, default 3); here a key
    sim["ndim"] = r.get_integer(SIMULATION, "ndim", 3);
    if (const int nd = std::get<int>(sim["ndim"]); nd < 1 || nd > 3)
        throw std::invalid_argument("The specified dimension (" + std::to_string(nd) + ") must be 1, 2 or 3!");

    // ---- flags (:37-47) ----
    const bool bosonic = r.get_boolean(SIMULATION, "bosonic", false);
    sim["bosonic"] = bosonic;
    sim["fixcom"] = r.get_boolean(SIMULATION, "fixcom", true);
    sim["pbc"] = r.get_boolean(SIMULATION, "pbc", false);
    const bool nmthermostat = r.get_boolean(SIMULATION, "nmthermostat", false);
    sim["nmthermostat"] = nmthermostat;

    // ---- initial positions (:49-95) ----
    std::string pos_type, pos_spec;
    if (!parse_token_parentheses(r.get(SIMULATION, "initial_position", "random"), pos_type, pos_spec))
        throw std::invalid_argument("The coordinate initialization method format is invalid!");
    if (!in_list(pos_type, {"random", "xyz", "grid"}))
        throw std::invalid_argument("The specified coordinate initialization method (" + pos_type +
                                    ") is not supported!");
    if (pos_type == "xyz" && is_index_pattern(pos_spec)) {
        sim["init_pos_first_index"] = static_cast<int>(!file_exists(format_index(pos_spec, 0)));
        pos_type = "xyz_formatted";
    }
    sim["init_pos_type"] = pos_type;
    if (pos_type == "xyz") sim["init_pos_xyz_filename"] = pos_spec;
    else if (pos_type == "xyz_formatted") sim["init_pos_xyz_filename_format"] = pos_spec;

    // ---- initial velocities (:97-138) ----
    std::string vel_type, vel_spec;
    if (!parse_token_parentheses(r.get(SIMULATION, "initial_velocity", "random"), vel_type, vel_spec))
        throw std::invalid_argument("The velocity initialization method format is invalid!");
    if (!in_list(vel_type, {"random", "manual"}))
        throw std::invalid_argument("The specified velocity initialization method (" + vel_type +
                                    ") is not supported!");
    if (vel_type == "manual" && !vel_spec.empty() && is_index_pattern(vel_spec)) {
        sim["init_vel_first_index"] = static_cast<int>(!file_exists(format_index(vel_spec, 0)));
        vel_type = "manual_formatted";
    }
    sim["init_vel_type"] = vel_type;
    if (vel_type == "manual_formatted") sim["init_vel_manual_filename_format"] = vel_spec;

    // ---- propagator and thermostat (:140-177) ----
    const std::string propagator = r.get_string(SIMULATION, "propagator", "cartesian");
    if (bosonic && propagator == "normal_modes")
        throw std::invalid_argument("Normal modes propogation is currently not available for bosons!");
    sim["propagator_type"] = propagator;
    if (!in_list(propagator, {"cartesian", "normal_modes"}))
        throw std::invalid_argument("The specified time propagator (" + propagator + ") is not supported!");
    const std::string thermostat = r.get_string(SIMULATION, "thermostat", "error");
    if (thermostat == "error") throw std::invalid_argument("Thermostat must be specified!");
    if (nmthermostat && thermostat == "none")
        throw std::invalid_argument("nmthermostat cannot be used in nve ensemble!");
    if (r.get_integer(SIMULATION, "nchains", -1) != -1 && (thermostat == "none" || thermostat == "langevin"))
        throw std::invalid_argument("nchains can only be used with Nose-Hoover thermostats!");
    sim["thermostat_type"] = thermostat;
    if (!in_list(thermostat, {"langevin", "nose_hoover", "nose_hoover_np", "nose_hoover_np_dim", "none"}))
        throw std::invalid_argument("The specified thermostat (" + thermostat + ") is not supported!");

    // ---- system (:179-200) ----
    sys["temperature"] = quantity("temperature", r.get(SYSTEM, "temperature", "1.0 kelvin"));
    if (const double t = std::get<double>(sys["temperature"]); t <= 0.0)
        throw std::invalid_argument("The specified temperature (" + fmt3(t) + " kelvin) is unphysical!");
    sys["natoms"] = r.get_integer(SYSTEM, "natoms", 1);
    if (const int n = std::get<int>(sys["natoms"]); n < 1)
        throw std::invalid_argument("The specified number of particles (" + std::to_string(n) +
                                    ") is smaller than one!");
    sys["mass"] = quantity("mass", r.get(SYSTEM, "mass", "1.0 dalton"));
    if (const double m = std::get<double>(sys["mass"]); m <= 0.0)
        throw std::invalid_argument("The provided mass (" + fmt3(m) + ") is unphysical!");
    sys["size"] = quantity("length", r.get(SYSTEM, "size", "1.0 picometer"));
    if (const double L = std::get<double>(sys["size"]); L <= 0.0)
        throw std::invalid_argument("The provided system size (" + fmt3(L) + ") is unphysical!");

    // ---- interaction potential (:205-227) ----
    const std::string int_name = r.get_string(INT_POTENTIAL, "name", "free");
    if (!in_list(int_name, {"aziz", "free", "harmonic", "dipole"}))
        throw std::invalid_argument("The specified interaction potential (" + int_name + ") is not supported!");
    interaction_pot["name"] = int_name;
    interaction_pot["cutoff"] = quantity("length", r.get(INT_POTENTIAL, "cutoff", "-1.0 angstrom"));
    if (int_name == "free") {
        interaction_pot["cutoff"] = 0.0;
    } else if (int_name == "harmonic") {
        interaction_pot["omega"] = quantity("energy", r.get(INT_POTENTIAL, "omega", "1.0 millielectronvolt"));
    } else if (int_name == "dipole") {
        interaction_pot["strength"] = r.get_real(INT_POTENTIAL, "strength", 1.0);
    }

    // ---- external potential (:229-252) ----
    const std::string ext_name = r.get_string(EXT_POTENTIAL, "name", "free");
    if (!in_list(ext_name, {"free", "harmonic", "double_well", "cosine"}))
        throw std::invalid_argument("The specified external potential (" + ext_name + ") is not supported!");
    external_pot["name"] = ext_name;
    if (ext_name == "harmonic") {
        external_pot["omega"] = quantity("energy", r.get(EXT_POTENTIAL, "omega", "1.0 millielectronvolt"));
    } else if (ext_name == "double_well") {
        external_pot["strength"] = quantity("energy", r.get(EXT_POTENTIAL, "strength", "1.0 millielectronvolt"));
        external_pot["location"] = quantity("length", r.get(EXT_POTENTIAL, "location", "1.0 angstrom"));
    } else if (ext_name == "cosine") {
        external_pot["amplitude"] = quantity("energy", r.get(EXT_POTENTIAL, "amplitude", "1.0 millielectronvolt"));
        external_pot["phase"] = r.get_real(EXT_POTENTIAL, "phase", 1.0);
    }

    // ---- output and observables (:254-265) ----
    states["positions"] = r.get(OUTPUT, "positions", "off");
    states["velocities"] = r.get(OUTPUT, "velocities", "off");
    states["forces"] = r.get(OUTPUT, "forces", "off");
    observables["energy"] = r.get(OBSERVABLES, "energy", "kelvin");
    observables["classical"] = r.get(OBSERVABLES, "classical", "off");
    observables["bosonic"] = r.get(OBSERVABLES, "bosonic", "off");
    observables["gsf"] = r.get(OBSERVABLES, "gsf", "off");
}

}  // namespace pimdb_host
