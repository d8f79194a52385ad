// kg_cli.hpp -- option parsing shared by the `index` and `search` command-line
// tools.  The reference parses the same options with boost::program_options
// (index.cpp:43-80, search.cpp:40-70); this is a small stand-alone parser for
// the forms those tools accept: --name value, --name=value, -X value, -Xvalue,
// boolean flags, and positional arguments in order.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace kgcli {

struct Args {
    std::map<std::string, std::string> opt;  // canonical long name -> value
    std::set<std::string> flags;
    std::vector<std::string> pos;
    bool has(std::string const &k) const { return opt.count(k) || flags.count(k); }
    template <typename T>
    T get(std::string const &k, T def) const {
        auto it = opt.find(k);
        if (it == opt.end()) return def;
        return convert<T>(it->second, k);
    }
    template <typename T>
    static T convert(std::string const &v, std::string const &k) {
        if constexpr (std::is_same<T, std::string>::value) {
            return v;
        } else {
            char *end = nullptr;
            T r;
            if (std::is_floating_point<T>::value)
                r = T(std::strtod(v.c_str(), &end));
            else if (std::is_signed<T>::value)
                r = T(std::strtol(v.c_str(), &end, 10));
            else
                r = T(std::strtoul(v.c_str(), &end, 10));
            if (v.empty() || *end) throw std::invalid_argument("bad value for --" + k + ": " + v);
            return r;
        }
    }
};

// valued: long name -> short alias ("" if none); boolean: long names.
inline Args parse(int argc, char **argv, std::map<std::string, std::string> const &valued,
                  std::set<std::string> const &boolean) {
    std::map<std::string, std::string> by_short;
    for (auto const &kv : valued)
        if (!kv.second.empty()) by_short[kv.second] = kv.first;
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        std::string name, value;
        bool have_value = false;
        if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
            name = s.substr(2);
            auto eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                have_value = true;
            }
        } else if (s.size() >= 2 && s[0] == '-' && !(s[1] >= '0' && s[1] <= '9')) {
            std::string sh = s.substr(1, 1);
            if (by_short.count(sh)) {
                name = by_short[sh];
                if (s.size() > 2) {
                    value = s.substr(2);
                    have_value = true;
                }
            } else if (sh == "h") {
                name = "help";
            } else if (sh == "v") {
                name = "version";
            } else {
                throw std::invalid_argument("unknown option " + s);
            }
        } else {
            a.pos.push_back(s);
            continue;
        }
        if (boolean.count(name)) {
            a.flags.insert(name);
        } else if (valued.count(name)) {
            if (!have_value) {
                if (i + 1 >= argc) throw std::invalid_argument("missing value for --" + name);
                value = argv[++i];
            }
            a.opt[name] = value;
        } else {
            throw std::invalid_argument("unknown option --" + name);
        }
    }
    return a;
}

}  // namespace kgcli
