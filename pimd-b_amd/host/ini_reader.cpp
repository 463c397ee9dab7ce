// ini_reader.cpp -- see ini_reader.hpp.
#include "ini_reader.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace pimdb_host {
namespace {

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

// Position of the first of `chars`, or of an inline ';' comment (a ';'
// preceded by whitespace), or npos.
size_t find_or_comment(const std::string& s, size_t from, const char* chars) {
    bool was_space = false;
    for (size_t i = from; i < s.size(); ++i) {
        const char c = s[i];
        if (chars && std::strchr(chars, c)) return i;
        if (was_space && c == ';') return i;
        was_space = std::isspace(static_cast<unsigned char>(c)) != 0;
    }
    return std::string::npos;
}

}  // namespace

IniReader::IniReader(const std::string& filename) {
    std::ifstream in(filename);
    if (!in.is_open()) {
        error_ = -1;
        return;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    parse(ss.str());
}

IniReader IniReader::from_string(const std::string& text) {
    IniReader r;
    r.parse(text);
    return r;
}

std::string IniReader::key(const std::string& section, const std::string& name) {
    std::string k = section + "=" + name;
    std::transform(k.begin(), k.end(), k.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return k;
}

void IniReader::parse(const std::string& text_in) {
    std::string text = text_in;
    if (text.size() >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB &&
        (unsigned char)text[2] == 0xBF)
        text.erase(0, 3);   // UTF-8 byte order mark
    std::istringstream in(text);
    std::string raw, section, prev_name;
    int lineno = 0;
    auto add = [&](const std::string& name, const std::string& value) {
        std::string& slot = values_[key(section, name)];
        if (!slot.empty()) slot += "\n";
        slot += value;
    };
    while (std::getline(in, raw)) {
        ++lineno;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        const bool indented = !raw.empty() && std::isspace(static_cast<unsigned char>(raw[0]));
        const std::string line = trim(raw);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (indented && !prev_name.empty()) {
            // continuation of the previous value
            const size_t c = find_or_comment(line, 0, nullptr);
            add(prev_name, trim(line.substr(0, c)));
            continue;
        }
        if (line[0] == '[') {
            const size_t close = find_or_comment(line, 1, "]");
            if (close != std::string::npos && line[close] == ']') {
                section = line.substr(1, close - 1);
                prev_name.clear();
            } else if (!error_) {
                error_ = lineno;
            }
            continue;
        }
        const size_t eq = find_or_comment(line, 0, "=:");
        if (eq != std::string::npos && (line[eq] == '=' || line[eq] == ':')) {
            const std::string name = trim(line.substr(0, eq));
            std::string rest = line.substr(eq + 1);
            const size_t c = find_or_comment(rest, 0, nullptr);
            if (c != std::string::npos) rest = rest.substr(0, c);
            add(name, trim(rest));
            prev_name = name;
        } else if (!error_) {
            error_ = lineno;
        }
    }
}

std::string IniReader::get(const std::string& section, const std::string& name, const std::string& dflt) const {
    const auto it = values_.find(key(section, name));
    return it == values_.end() ? dflt : it->second;
}

std::string IniReader::get_string(const std::string& section, const std::string& name, const std::string& dflt) const {
    const std::string s = get(section, name, "");
    return s.empty() ? dflt : s;
}

long IniReader::get_long(const std::string& section, const std::string& name, long dflt) const {
    const std::string s = get(section, name, "");
    char* end = nullptr;
    const long n = std::strtol(s.c_str(), &end, 0);
    return end > s.c_str() ? n : dflt;
}

int IniReader::get_integer(const std::string& section, const std::string& name, int dflt) const {
    const std::string s = get(section, name, "");
    return s.empty() ? dflt : std::atoi(s.c_str());
}

double IniReader::get_real(const std::string& section, const std::string& name, double dflt) const {
    const std::string s = get(section, name, "");
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    return end > s.c_str() ? v : dflt;
}

bool IniReader::get_boolean(const std::string& section, const std::string& name, bool dflt) const {
    std::string s = get(section, name, "");
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (s == "true" || s == "yes" || s == "on" || s == "1") return true;
    if (s == "false" || s == "no" || s == "off" || s == "0") return false;
    return dflt;
}

bool IniReader::has_value(const std::string& section, const std::string& name) const {
    return values_.count(key(section, name)) != 0;
}

}  // namespace pimdb_host
