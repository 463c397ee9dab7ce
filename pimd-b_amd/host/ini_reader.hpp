// ini_reader.hpp -- the configuration-file reader the reference's Params uses
// (INIReader over inih: libs/inireader.cpp, libs/ini.c), with the same
// getter semantics: keys are case-insensitive "section=name"; a repeated key
// or a continuation line (leading whitespace) appends "\n" + value; start-of-
// line comments begin with ';' or '#', inline comments with whitespace + ';'.
#pragma once

#include <map>
#include <string>

namespace pimdb_host {

class IniReader {
public:
    IniReader() = default;
    // Reads a file; parse_error() is 0 on success, -1 when the file cannot be
    // opened, or the first line number with a syntax error.
    explicit IniReader(const std::string& filename);
    static IniReader from_string(const std::string& text);

    int parse_error() const { return error_; }
    std::string get(const std::string& section, const std::string& name, const std::string& dflt) const;
    std::string get_string(const std::string& section, const std::string& name, const std::string& dflt) const;
    long get_long(const std::string& section, const std::string& name, long dflt) const;
    int get_integer(const std::string& section, const std::string& name, int dflt) const;
    double get_real(const std::string& section, const std::string& name, double dflt) const;
    bool get_boolean(const std::string& section, const std::string& name, bool dflt) const;
    bool has_value(const std::string& section, const std::string& name) const;

private:
    void parse(const std::string& text);
    static std::string key(const std::string& section, const std::string& name);
    std::map<std::string, std::string> values_;
    int error_ = 0;
};

}  // namespace pimdb_host
