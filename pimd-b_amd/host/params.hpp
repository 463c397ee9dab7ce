// params.hpp -- the run configuration read from config.ini, with the
// reference's sections, keys, defaults, unit handling and validation
// messages (Params::Params, src/params.cpp:8-258).
//
// The maps keep the reference's names and value types so a caller reads
// them as it reads Params: sim, sys, interaction_pot, external_pot, states,
// observables.
#pragma once

#include <map>
#include <string>
#include <variant>

#include "ini_reader.hpp"

namespace pimdb_host {

using Value = std::variant<int, long, unsigned int, double, bool, std::string>;
using VariantMap = std::map<std::string, Value>;
using StringMap = std::map<std::string, std::string>;

namespace sections {
const std::string SYSTEM = "system";
const std::string SIMULATION = "simulation";
const std::string EXT_POTENTIAL = "external_potential";
const std::string INT_POTENTIAL = "interaction_potential";
const std::string OUTPUT = "output";
const std::string OBSERVABLES = "observables";
}  // namespace sections

class Params {
public:
    // Throws std::invalid_argument with the reference's messages.
    explicit Params(const std::string& filename);
    static Params from_string(const std::string& text);

    IniReader reader;
    VariantMap sim, sys, interaction_pot, external_pot;
    StringMap states, observables;

    // "token(value)", "token()" or "token" (params.cpp:312-330)
    static bool parse_token_parentheses(const std::string& input, std::string& token, std::string& value);

    template <class T>
    T get(const VariantMap& m, const std::string& k) const {
        return std::get<T>(m.at(k));
    }

private:
    Params() = default;
    void load();
};

// Formats "{}" / "{:0N}" / "{:N}" in a file-name pattern with an integer (the
// reference formats per-bead file names with std::vformat).
std::string format_index(const std::string& pattern, int value);
bool is_index_pattern(const std::string& pattern);

}  // namespace pimdb_host
