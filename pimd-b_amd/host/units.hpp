// units.hpp -- unit conversion of input quantities ("<number> <unit>") to
// atomic units and back, with metric prefixes.
//
// Reference: Units::convertToInternal / convertToUser and the unit tables
// (src/units.cpp:17-64, include/units.h:18-153).  The conversion factors are
// the reference's (i-PI's low-precision values) and are used verbatim: every
// number the simulation reports depends on them.
#pragma once

#include <string>
#include <utility>

namespace pimdb_host {

// kB = hbar = 1; amu in electron masses (include/units.h:18-22)
constexpr double AMU = 1822.8885;

// Factor that takes `unit` (with an optional metric prefix) of `family` to
// atomic units.  Throws std::invalid_argument for an unknown family or unit,
// with the reference's messages.
double unit_factor(const std::string& family, const std::string& unit);
double to_internal(const std::string& family, const std::string& unit, double number);
double to_user(const std::string& family, const std::string& unit, double number);

// "<number> <unit>" -> (number, unit); throws std::invalid_argument
// ("Invalid input format") when either is missing (params.cpp:281-293).
std::pair<double, std::string> parse_quantity(const std::string& text);
// parse_quantity followed by to_internal (Params::getQuantity, params.cpp:303-309).
double quantity(const std::string& family, const std::string& text);

}  // namespace pimdb_host
