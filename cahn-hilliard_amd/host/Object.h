// Object.h -- configuration helpers and logging of the host-side drop-in.
//
// Mirrors src/Object.h:22-112 of the reference: _config_value /
// _config_optional_value / _config_array_values on dotted TOML paths (a scalar
// is broadcast to output_size entries), info / warning / critical where
// critical logs and calls exit(1).  Integers given as integral floats (e.g.
// `steps = 1e7`) are accepted for integer options, ints for double options.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <string_view>
#include <type_traits>
#include <vector>

#include "toml_lite.h"

#define SQR(X) ((X) * (X))
#define CUB(X) ((X) * (X) * (X))

namespace ch {

namespace detail {
inline void format_into(std::ostringstream &os, const char *f) { os << f; }
template <typename A, typename... R>
void format_into(std::ostringstream &os, const char *f, const A &a, const R &...r) {
    while(*f) {
        if(f[0] == '{' && f[1] == '}') {
            os << a;
            format_into(os, f + 2, r...);
            return;
        }
        os << *f++;
    }
}
}  // namespace detail

// "{}" placeholders replaced by the arguments, in order
template <typename... A>
std::string format(const char *f, const A &...a) {
    std::ostringstream os;
    detail::format_into(os, f, a...);
    return os.str();
}

class Object {
public:
    virtual ~Object() = default;
    virtual std::string_view name() const = 0;

    template <typename... A>
    void info(const char *f, const A &...a) const {
        std::fprintf(stderr, "[info] %s (source: %.*s)\n", format(f, a...).c_str(), (int)name().size(), name().data());
    }
    template <typename... A>
    void warning(const char *f, const A &...a) const {
        std::fprintf(stderr, "[warning] %s (source: %.*s)\n", format(f, a...).c_str(), (int)name().size(), name().data());
    }
    template <typename... A>
    [[noreturn]] void critical(const char *f, const A &...a) const {
        std::fprintf(stderr, "[critical] %s (error source: %.*s)\n", format(f, a...).c_str(), (int)name().size(),
                     name().data());
        std::exit(1);
    }

    const chtoml::Node *_config_node(const chtoml::Table &t, const std::string &path, bool mandatory) const {
        const chtoml::Node *n = chtoml::find(t, path);
        if(mandatory && !n) critical("Mandatory option '{}' not found", path);
        return n;
    }

    template <typename T>
    bool _convert(const chtoml::Node &n, T &out) const {
        if constexpr(std::is_same_v<T, bool>) {
            if(n.kind != chtoml::Node::BOOL) return false;
            out = n.b;
        }
        else if constexpr(std::is_integral_v<T>) {
            if(n.kind == chtoml::Node::INT) out = (T)n.i;
            else if(n.kind == chtoml::Node::FLOAT && std::floor(n.f) == n.f) out = (T)n.f;
            else return false;
        }
        else if constexpr(std::is_floating_point_v<T>) {
            if(!n.is_number()) return false;
            out = (T)n.as_double();
        }
        else {
            if(n.kind != chtoml::Node::STRING) return false;
            out = n.s;
        }
        return true;
    }

    template <typename T>
    T _config_value(const chtoml::Table &t, const std::string &path) const {
        const chtoml::Node *n = _config_node(t, path, true);
        T v{};
        if(!_convert(*n, v)) critical("Option '{}' has the wrong type", path);
        return v;
    }

    template <typename T>
    T _config_optional_value(const chtoml::Table &t, const std::string &path, T default_value) const {
        const chtoml::Node *n = _config_node(t, path, false);
        T v{};
        if(!n || !_convert(*n, v)) return default_value;
        return v;
    }

    template <typename T>
    std::vector<T> _config_array_values(const chtoml::Table &t, const std::string &path, size_t output_size = 0) const {
        const chtoml::Node *n = _config_node(t, path, true);
        std::vector<T> out;
        T v{};
        if(n->kind != chtoml::Node::ARRAY && _convert(*n, v) && output_size > 0) {
            out.assign(output_size, v);
        }
        else if(n->kind == chtoml::Node::ARRAY) {
            for(auto &e : n->arr) {
                if(!_convert(e, v)) critical("Option '{}': cannot convert an element", path);
                out.push_back(v);
            }
        }
        if(output_size > 0 && out.size() != output_size)
            critical("Option '{}' should be a single value or an array of size {}, not {}", path, output_size, out.size());
        return out;
    }
};

}  // namespace ch
