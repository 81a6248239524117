// .ini-driven parameters: ConfigMap and the ADVParams / Conv1dParams structs.
//
// Same names, keys, defaults and parsing behaviour as the reference's
// src/config/ (ConfigMap.cpp:35-80, INIReader.cpp:43-100, ini.cpp:25-146,
// AdvectionParams.cpp:64-131, Conv1dParams.cpp:44-76), written for this
// library.  Parsing rules kept because they change parameter values:
//  - keys are "section.name", lower-cased (case-insensitive lookups);
//  - ';' starts an inline comment only after whitespace, '#' only at the
//    start of a line (so "n0 = 8 # c" keeps "8 # c" and strtol still reads 8,
//    but a bool "true # c" is not recognised and falls back to the default);
//  - an indented non-empty line continues (replaces) the previous value;
//  - lines are read in chunks of at most 199 characters, section and key
//    names are truncated to 49 characters;
//  - getInteger uses strtol(..., 0) (0x.. hex, leading 0 octal);
//  - getFloat parses with strtof, so a value arrives rounded to float and
//    then widened to double (dt = 0.001 becomes 0.0010000000474974513).
#pragma once

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "bkma_amd.h"

namespace bkma {

class INIReader {
  public:
    explicit INIReader(const std::string &filename) { _error = parse_file(filename); }
    INIReader(char *&buffer, int buffer_size) { _error = parse_buffer(std::string(buffer, buffer_size)); }
    virtual ~INIReader() = default;

    // 0 on success, -1 if the file cannot be opened, else the first bad line.
    int ParseError() const { return _error; }

    std::string getString(const std::string &section, const std::string &name,
                          const std::string &default_value) const {
        auto it = _values.find(makeKey(section, name));
        return it == _values.end() ? default_value : it->second;
    }
    void setString(const std::string &section, const std::string &name, const std::string &value) {
        _values[makeKey(section, name)] = value;
    }
    long getInteger(const std::string &section, const std::string &name, long default_value) const {
        const std::string s = getString(section, name, "");
        const char *v = s.c_str();
        char *end;
        const long n = std::strtol(v, &end, 0);
        return end > v ? n : default_value;
    }
    void setInteger(const std::string &section, const std::string &name, long value) {
        setString(section, name, std::to_string(value));
    }
    const std::map<std::string, std::string> &values() const { return _values; }

    static std::string makeKey(const std::string &section, const std::string &name) {
        std::string key = section + "." + name;
        for (auto &c : key) c = (char)std::tolower((unsigned char)c);
        return key;
    }

  private:
    static constexpr size_t kMaxLine = 200, kMaxSection = 50, kMaxName = 50;
    std::map<std::string, std::string> _values;
    int _error = 0;

    static std::string rstrip(const std::string &s) {
        size_t e = s.size();
        while (e > 0 && std::isspace((unsigned char)s[e - 1])) --e;
        return s.substr(0, e);
    }
    static size_t lskip(const std::string &s, size_t i = 0) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        return i;
    }
    // Position of c, or of a ';' preceded by whitespace, or npos.
    static size_t find_char_or_comment(const std::string &s, size_t from, char c) {
        bool was_ws = false;
        for (size_t i = from; i < s.size(); ++i) {
            if (s[i] == c || (was_ws && s[i] == ';')) return i;
            was_ws = std::isspace((unsigned char)s[i]) != 0;
        }
        return std::string::npos;
    }
    static std::string trunc(const std::string &s, size_t cap) { return s.size() < cap ? s : s.substr(0, cap - 1); }

    int parse_line(const std::string &raw, std::string &section, std::string &prev_name, int lineno, int error) {
        const std::string line = rstrip(raw);
        const size_t st = lskip(line);
        if (!prev_name.empty() && st < line.size() && st > 0) {
            _values[makeKey(section, prev_name)] = line.substr(st);  // continuation replaces the value
            return error;
        }
        if (st >= line.size()) return error;
        const char c0 = line[st];
        if (c0 == ';' || c0 == '#') return error;
        if (c0 == '[') {
            const size_t e = find_char_or_comment(line, st + 1, ']');
            if (e != std::string::npos && line[e] == ']') {
                section = trunc(line.substr(st + 1, e - st - 1), kMaxSection);
                prev_name.clear();
            } else if (!error) {
                error = lineno;
            }
            return error;
        }
        const size_t e = find_char_or_comment(line, st, '=');
        if (e != std::string::npos && line[e] == '=') {
            const std::string name = rstrip(line.substr(st, e - st));
            std::string value = line.substr(lskip(line, e + 1));
            const size_t cm = find_char_or_comment(value, 0, '\0');
            if (cm != std::string::npos && value[cm] == ';') value = value.substr(0, cm);
            value = rstrip(value);
            prev_name = trunc(name, kMaxName);
            _values[makeKey(section, name)] = value;
        } else if (!error) {
            error = lineno;
        }
        return error;
    }

    int parse_file(const std::string &filename) {
        FILE *f = std::fopen(filename.c_str(), "r");
        if (!f) return -1;
        char buf[kMaxLine];
        std::string section, prev_name;
        int lineno = 0, error = 0;
        while (std::fgets(buf, sizeof(buf), f)) error = parse_line(buf, section, prev_name, ++lineno, error);
        std::fclose(f);
        return error;
    }

    int parse_buffer(const std::string &text) {
        std::istringstream in(text);
        std::string section, prev_name, line;
        int lineno = 0, error = 0;
        while (std::getline(in, line)) error = parse_line(line, section, prev_name, ++lineno, error);
        return error;
    }
};

class ConfigMap : public INIReader {
  public:
    explicit ConfigMap(const std::string &filename) : INIReader(filename) {}
    ConfigMap(char *&buffer, int buffer_size) : INIReader(buffer, buffer_size) {}

    double getFloat(const std::string &section, const std::string &name, float default_value) const {
        const std::string s = getString(section, name, "");
        const char *v = s.c_str();
        char *end;
        const float f = std::strtof(v, &end);
        return end > v ? (double)f : (double)default_value;
    }
    void setFloat(const std::string &section, const std::string &name, float value) {
        std::stringstream ss;
        ss << value;
        setString(section, name, ss.str());
    }
    bool getBool(const std::string &section, const std::string &name, bool default_value) const {
        bool val = default_value;
        const std::string s = getString(section, name, "");
        if (s == "1" || s == "yes" || s == "true" || s == "on") val = true;
        if (s == "0" || s == "no" || s == "false" || s == "off") val = false;
        if (s.empty()) val = default_value;
        return val;
    }
    void setBool(const std::string &section, const std::string &name, bool value) {
        setString(section, name, value ? "true" : "false");
    }
};

// Single-process build: the file is read directly (the reference's MPI
// broadcast, ConfigMap.cpp:96-147, is compiled out there as well).
inline ConfigMap broadcast_parameters(const std::string &filename) { return ConfigMap(filename); }

// ---------------------------------------------------------------------------
// ADVParams (src/config/AdvectionParams.hpp:10-57): the POD part is exactly
// bkma_adv_params, so it crosses the C ABI unchanged.
// ---------------------------------------------------------------------------
struct ADVParamsNonCopyable;

struct ADVParams : bkma_adv_params {
    ADVParams() {
        bkma_adv_params &b = *this;
        std::memset(&b, 0, sizeof(b));
        gpu = 0;
        percent_loc = 1.0f;
        outputSolution = 0;
        maxIter = 100;
        n1 = 1024;
        n0 = 32;
        n2 = 32;
        pref_wg_size = 128;
        seq_size0 = seq_size2 = 1;
        dt = 0.0001;
        minRealX = 0;
        maxRealX = 1;
        realWidthX = 1;
        minRealVx = 0;
        maxRealVx = 1;
        update_deltas();
    }
    inline explicit ADVParams(const ADVParamsNonCopyable &other);
    void update_deltas() { bkma_adv_update_deltas(this); }
};

struct ADVParamsNonCopyable : ADVParams {
    std::string kernelImpl;
    bool inplace = true;

    ADVParamsNonCopyable() = default;
    explicit ADVParamsNonCopyable(const ADVParams &other) : ADVParams(other) {}

    // AdvectionParams.cpp:64-92
    void setup(const ConfigMap &c) {
        n1 = c.getInteger("problem", "n1", 1024);
        n0 = c.getInteger("problem", "n0", 1024);
        n2 = c.getInteger("problem", "n2", 1024);
        maxIter = c.getInteger("problem", "maxIter", 50);
        dt = c.getFloat("problem", "dt", 0.0001f);
        minRealX = c.getFloat("problem", "minRealX", 0.0f);
        maxRealX = c.getFloat("problem", "maxRealX", 1.0f);
        minRealVx = c.getFloat("problem", "minRealVx", -1.0f);
        maxRealVx = c.getFloat("problem", "maxRealVx", 1.0f);
        kernelImpl = c.getString("impl", "kernelImpl", "AdaptiveWg");
        inplace = c.getBool("impl", "inplace", true);
        gpu = c.getBool("optimization", "gpu", true);
        pref_wg_size = c.getInteger("optimization", "pref_wg_size", 512);
        seq_size0 = c.getInteger("optimization", "seq_size0", 1);
        seq_size2 = c.getInteger("optimization", "seq_size2", 1);
        outputSolution = c.getBool("io", "outputSolution", false);
        update_deltas();
    }

    // AdvectionParams.cpp:107-131
    void print(std::ostream &os = std::cout) const {
        os << "##########################\nRuntime parameters:\n##########################\n";
        os << "kernelImpl  : " << kernelImpl << "\n";
        os << "inplace     : " << inplace << "\n";
        os << "gpu         : " << gpu << "\n";
        os << "maxIter     : " << maxIter << "\n";
        os << "n0 (nvx)    : " << n0 << "\n";
        os << "n1 (nx)     : " << n1 << "\n";
        os << "n2          : " << n2 << "\n";
        os << "pref_wg_size: " << pref_wg_size << "\n";
        os << "seq_size0   : " << seq_size0 << "\n";
        os << "seq_size2   : " << seq_size2 << "\n";
        os << "dt          : " << dt << "\n";
        os << "dx          : " << dx << "\n";
        os << "dvx         : " << dvx << "\n";
        os << "minRealX    : " << minRealX << "\n";
        os << "maxRealX    : " << maxRealX << "\n";
        os << "minRealVx   : " << minRealVx << "\n";
        os << "maxRealVx   : " << maxRealVx << "\n\n";
    }
};

inline ADVParams::ADVParams(const ADVParamsNonCopyable &other) {
    const bkma_adv_params &src = other;
    bkma_adv_params &dst = *this;
    dst = src;
}

// ---------------------------------------------------------------------------
// Conv1dParams (src/config/Conv1dParams.hpp:10-44, Conv1dParams.cpp:44-76)
// ---------------------------------------------------------------------------
struct Conv1dParams {
    size_t length = 1024;
    unsigned short channel_in = 3;
    unsigned short channel_out = 3;  // = channel_in (constraint)
    unsigned short k = 3;
    size_t total_batch_size = 262144;
    size_t batch_size_n2 = 512;
    size_t n0 = 0, n1 = 0, n2 = 0, n_write = 0;
    bool gpu = true;
    unsigned short pref_wg_size = 512;
    size_t seq_size0 = 1, seq_size2 = 1;
    bool inplace = true;

    static size_t compute_output_size(size_t Lin, unsigned short kernel_size) { return Lin - (kernel_size - 1); }
};

struct Conv1dParamsNonCopyable : Conv1dParams {
    std::string kernelImpl;

    void setup(const ConfigMap &c) {
        length = c.getInteger("problem", "length", 1024);
        channel_in = (unsigned short)c.getInteger("problem", "channel_in", 3);
        channel_out = channel_in;
        k = (unsigned short)c.getInteger("problem", "k", 3);
        total_batch_size = c.getInteger("problem", "total_batch_size", 262144);
        batch_size_n2 = c.getInteger("problem", "batch_size_n2", 512);
        n0 = total_batch_size / batch_size_n2;
        n1 = length * channel_out;
        n2 = batch_size_n2;
        n_write = compute_output_size(length, k);
        kernelImpl = c.getString("impl", "kernelImpl", "AdaptiveWg");
        inplace = c.getBool("impl", "inplace", true);
        gpu = c.getBool("optimization", "gpu", true);
        pref_wg_size = (unsigned short)c.getInteger("optimization", "pref_wg_size", 512);
        seq_size0 = c.getInteger("optimization", "seq_size0", 1);
        seq_size2 = c.getInteger("optimization", "seq_size2", 1);
    }

    void print(std::ostream &os = std::cout) const {
        os << "##########################\nRuntime parameters:\n##########################\n";
        os << "kernelImpl   : " << kernelImpl << "\n";
        os << "inplace      : " << inplace << "\n";
        os << "gpu          : " << gpu << "\n";
        os << "n0           : " << n0 << "\n";
        os << "n1           : " << n1 << "\n";
        os << "n2           : " << n2 << "\n";
        os << "pref_wg_size : " << pref_wg_size << "\n";
        os << "seq_size0    : " << seq_size0 << "\n";
        os << "seq_size2    : " << seq_size2 << "\n";
        os << "batch_size   : " << total_batch_size << "\n";
        os << "length       : " << length << "\n";
        os << "channels(i/o): " << channel_in << "\n";
        os << "k            : " << k << "\n";
        os << "batch_n2     : " << batch_size_n2 << "\n";
        os << "n_write      : " << n_write << "\n\n";
    }
};

}  // namespace bkma
