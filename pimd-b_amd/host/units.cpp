// units.cpp -- see units.hpp.
#include "units.hpp"

#include <cstring>
#include <sstream>
#include <stdexcept>

namespace pimdb_host {
namespace {

struct UnitRow {
    const char* family;
    const char* unit;
    double factor;
};

// include/units.h:51-153.  Every family also accepts "", "automatic" and
// "atomic_unit" (factor 1), handled in find_factor.
const UnitRow kUnits[] = {
    {"energy", "electronvolt", 0.036749326},
    {"energy", "j/mol", 0.00000038087989},
    {"energy", "cal/mol", 0.0000015946679},
    {"energy", "kelvin", 3.1668152e-06},
    {"temperature", "kelvin", 3.1668152e-06},
    {"time", "second", 4.1341373e16},
    {"frequency", "inversecm", 4.5563353e-06},
    {"frequency", "hertz*rad", 2.4188843e-17},
    {"frequency", "hertz", 1.5198298e-16},
    {"length", "angstrom", 1.8897261},
    {"length", "meter", 1.8897261e10},
    {"length", "radian", 1.00},
    {"length", "degree", 0.017453292519943295},
    {"volume", "angstrom3", 6.748334231},
    {"velocity", "angstrom/ps", 4.5710289e-5},
    {"velocity", "m/s", 4.5710289e-7},
    {"mass", "dalton", 1.00 * AMU},
    {"mass", "amu", 1.00 * AMU},
    {"mass", "electronmass", 1.00},
    {"pressure", "bar", 3.398827377e-9},
    {"pressure", "atmosphere", 3.44386184e-9},
    {"pressure", "pascal", 3.398827377e-14},
    {"pressure", "ev/ang3", 0.0054456877},
    {"density", "g/cm3", 162.67263},
    {"force", "newton", 12137805},
    {"force", "ev/ang", 0.019446904},
    {"hessian", "ev/ang^2", 0.010290858},
};

const char* const kFamilies[] = {"undefined", "energy", "temperature", "time", "frequency", "ms-momentum",
                                 "length", "volume", "velocity", "momentum", "mass", "pressure", "density",
                                 "force", "hessian"};

struct Prefix {
    const char* name;
    double factor;
};
const Prefix kPrefixes[] = {
    {"yotta", 1e24}, {"zetta", 1e21}, {"exa", 1e18},   {"peta", 1e15},   {"tera", 1e12},
    {"giga", 1e9},   {"mega", 1e6},   {"kilo", 1e3},   {"hecto", 1e2},   {"deci", 1e-1},
    {"centi", 1e-2}, {"milli", 1e-3}, {"micro", 1e-6}, {"nano", 1e-9},   {"pico", 1e-12},
    {"femto", 1e-15}, {"atto", 1e-18}, {"zepto", 1e-21}, {"yocto", 1e-24},
};

bool known_family(const std::string& f) {
    for (const char* k : kFamilies)
        if (f == k) return true;
    return false;
}

// Leading prefixes are consumed greedily, the last one counts; the rest is
// the base unit (the reference's "(p1|p2|...)*(.*)" match, units.cpp:17-40).
void split_prefix(const std::string& unit, double& prefix, std::string& base) {
    prefix = 1.0;
    size_t pos = 0;
    for (bool again = true; again;) {
        again = false;
        for (const Prefix& p : kPrefixes) {
            const size_t n = std::strlen(p.name);
            if (unit.compare(pos, n, p.name) == 0) {
                prefix = p.factor;
                pos += n;
                again = true;
                break;
            }
        }
    }
    base = unit.substr(pos);
}

}  // namespace

// Base factor and prefix factor of a unit; number * base * prefix is the
// reference's evaluation order.
static void factors(const std::string& family, const std::string& unit, double& base_f, double& prefix) {
    if (!known_family(family)) throw std::invalid_argument(family + " is an undefined units kind.");
    std::string base;
    split_prefix(unit, prefix, base);
    if (base.empty() || base == "automatic" || base == "atomic_unit") {
        base_f = 1.00;
        return;
    }
    for (const UnitRow& r : kUnits) {
        if (family == r.family && base == r.unit) {
            base_f = r.factor;
            return;
        }
    }
    throw std::invalid_argument(base + " is an undefined unit for kind " + family + ".");
}

double unit_factor(const std::string& family, const std::string& unit) { return to_internal(family, unit, 1.0); }

double to_internal(const std::string& family, const std::string& unit, double number) {
    if (family == "number") return number;
    double base_f, prefix;
    factors(family, unit, base_f, prefix);
    return number * base_f * prefix;
}

double to_user(const std::string& family, const std::string& unit, double number) {
    return number / to_internal(family, unit, 1.0);
}

std::pair<double, std::string> parse_quantity(const std::string& text) {
    std::istringstream in(text);
    double value;
    std::string unit;
    if (in >> value >> std::ws >> unit) return {value, unit};
    throw std::invalid_argument("Invalid input format");
}

double quantity(const std::string& family, const std::string& text) {
    const auto q = parse_quantity(text);
    return to_internal(family, q.second, q.first);
}

}  // namespace pimdb_host
