// Command-line parsing and console tables for the per-technique executables.
//
// The executables accept the same flags as the reference's cxxopts (v2.2.1)
// declarations, e.g. ip.cu:37-43: "--name value" or "--name=value"; a bool
// flag given without a value is true ("--warp"), "--warp=false" clears it.
// Parse errors print the message and exit with status 1, as the reference's
// catch of cxxopts::OptionException does (ip.cu:200-203).
// The box tables reproduce the reference's fmt layouts ("┌{0:─^{1}}┐",
// "│{3: ^{2}}|{4: ^{2}}│", ip.cu:16-19, 59-68).
#pragma once

#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace si {

class CliError : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

class Options {
   public:
    Options(std::string program, std::string description)
        : program_(std::move(program)), description_(std::move(description)) {}

    Options& add(const std::string& name, const std::string& help, std::string* target) {
        return addImpl(name, help, Kind::String, target);
    }
    Options& add(const std::string& name, const std::string& help, uint32_t* target) {
        return addImpl(name, help, Kind::U32, target);
    }
    Options& add(const std::string& name, const std::string& help, uint64_t* target) {
        return addImpl(name, help, Kind::U64, target);
    }
    Options& add(const std::string& name, const std::string& help, double* target) {
        return addImpl(name, help, Kind::Double, target);
    }
    Options& add(const std::string& name, const std::string& help, bool* target) {
        return addImpl(name, help, Kind::Bool, target);
    }
    Options& addFlag(const std::string& name, const std::string& help) {
        return addImpl(name, help, Kind::Flag, nullptr);
    }

    // Parses argv; positional arguments are returned in order.
    std::vector<std::string> parse(int argc, char** argv) {
        std::vector<std::string> positional;
        for (int i = 1; i < argc; ++i) {
            std::string arg = argv[i];
            if (arg.size() < 3 || arg[0] != '-' || arg[1] != '-') {
                positional.push_back(arg);
                continue;
            }
            std::string name = arg.substr(2), value;
            bool hasValue = false;
            auto eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                hasValue = true;
            }
            auto it = opts_.find(name);
            if (it == opts_.end()) throw CliError("Option '" + name + "' does not exist");
            Opt& o = it->second;
            if (o.kind == Kind::Bool || o.kind == Kind::Flag) {
                if (!hasValue) value = "true";
            } else if (!hasValue) {
                if (i + 1 >= argc) throw CliError("Option '" + name + "' is missing an argument");
                value = argv[++i];
            }
            store(name, o, value);
            ++count_[name];
        }
        return positional;
    }

    size_t count(const std::string& name) const {
        auto it = count_.find(name);
        return it == count_.end() ? 0 : it->second;
    }

    std::string help() const {
        std::string s = description_ + "\nUsage:\n  " + program_ + " [OPTION...]\n\n";
        for (auto& name : order_) {
            const Opt& o = opts_.at(name);
            std::string lhs = "      --" + name;
            if (o.kind != Kind::Bool && o.kind != Kind::Flag) lhs += " arg";
            if (lhs.size() < 30) lhs.resize(30, ' ');
            s += lhs + "  " + o.help + "\n";
        }
        return s;
    }

   private:
    enum class Kind { String, U32, U64, Double, Bool, Flag };
    struct Opt {
        std::string help;
        Kind kind;
        void* target;
    };
    Options& addImpl(const std::string& name, const std::string& help, Kind kind, void* target) {
        opts_[name] = Opt{help, kind, target};
        order_.push_back(name);
        return *this;
    }
    static uint64_t parseUnsigned(const std::string& name, const std::string& v, uint64_t max) {
        uint64_t x = 0;
        auto r = std::from_chars(v.data(), v.data() + v.size(), x);
        if (v.empty() || r.ec != std::errc() || r.ptr != v.data() + v.size() || x > max)
            throw CliError("Argument '" + v + "' failed to parse (option '" + name + "')");
        return x;
    }
    void store(const std::string& name, Opt& o, const std::string& v) {
        switch (o.kind) {
            case Kind::String: *static_cast<std::string*>(o.target) = v; break;
            case Kind::U32: *static_cast<uint32_t*>(o.target) = uint32_t(parseUnsigned(name, v, 0xFFFFFFFFull)); break;
            case Kind::U64: *static_cast<uint64_t*>(o.target) = parseUnsigned(name, v, ~0ull); break;
            case Kind::Double: {
                char* end = nullptr;
                double x = std::strtod(v.c_str(), &end);
                if (v.empty() || *end) throw CliError("Argument '" + v + "' failed to parse (option '" + name + "')");
                *static_cast<double*>(o.target) = x;
                break;
            }
            case Kind::Bool:
            case Kind::Flag: {
                bool b;
                if (v == "true" || v == "1" || v == "yes") b = true;
                else if (v == "false" || v == "0" || v == "no") b = false;
                else throw CliError("Argument '" + v + "' failed to parse (option '" + name + "')");
                if (o.target) *static_cast<bool*>(o.target) = b;
                break;
            }
        }
    }
    std::string program_, description_;
    std::map<std::string, Opt> opts_;
    std::vector<std::string> order_;
    std::map<std::string, size_t> count_;
};

// ------------------------------------------------------------------ tables --
// Number of code points in a UTF-8 string.
inline size_t u8len(const std::string& s) {
    size_t n = 0;
    for (unsigned char c : s) n += (c & 0xC0) != 0x80;
    return n;
}

// fmt-style centre alignment: left pad = floor(pad/2).
inline std::string center(const std::string& s, size_t width, const std::string& fill = " ") {
    size_t len = u8len(s);
    if (len >= width) return s;
    size_t pad = width - len, left = pad / 2, right = pad - left;
    std::string out;
    for (size_t i = 0; i < left; ++i) out += fill;
    out += s;
    for (size_t i = 0; i < right; ++i) out += fill;
    return out;
}

inline std::string toText(const std::string& s) { return s; }
inline std::string toText(const char* s) { return s; }
inline std::string toText(float v) {
    char b[32];
    return std::string(b, std::to_chars(b, b + 32, v).ptr);
}
inline std::string toText(double v) {
    char b[32];
    return std::string(b, std::to_chars(b, b + 32, v).ptr);
}
template <typename T>
inline std::string toText(T v) {
    return std::to_string(v);
}

// Title box:  ┌───┐ / │ title | / └───┘  (ip.cu:16-20)
inline void printTitle(const std::string& title) {
    std::printf("┌%s┐\n│%s|\n└%s┘\n", center("", 51, "─").c_str(), center(title, 51).c_str(),
                center("", 51, "─").c_str());
}

// Key/value box (ip.cu:59-68, 86-106).
inline void printTable(const std::string& title, const std::vector<std::pair<std::string, std::string>>& rows) {
    std::printf("┌%s┐\n", center(title, 51, "─").c_str());
    for (auto& r : rows) std::printf("│%s|%s│\n", center(r.first, 25).c_str(), center(r.second, 25).c_str());
    std::printf("└%s┘\n", center("", 51, "─").c_str());
}

}  // namespace si
