// toml_lite.h -- the TOML subset the cahn-hilliard input files use.
//
// Supported: `key = value` lines, [table] and [a.b] headers, dotted keys,
// basic "strings" and 'literal strings', integers (with _ separators),
// floats (1e7, 0.5, inf, nan), booleans, one-line arrays of scalars, inline
// tables { a = 1, b = "x" }, and # comments.  That covers every key the
// reference reads (README.md:24-53, examples/*.toml, tests/data/*.toml).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace chtoml {

struct Node;
using Table = std::map<std::string, Node>;

struct Node {
    enum Kind { NONE, INT, FLOAT, BOOL, STRING, ARRAY, TABLE } kind = NONE;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    std::string s;
    std::vector<Node> arr;
    std::shared_ptr<Table> tab;

    bool is_number() const { return kind == INT || kind == FLOAT; }
    double as_double() const { return kind == INT ? (double)i : f; }
};

// Parses `text`; on failure returns false and sets `err` ("line N: ...").
bool parse(const std::string &text, Table &out, std::string &err);
bool parse_file(const std::string &path, Table &out, std::string &err);

// Looks up a dotted path ("mobility.type"); nullptr when absent.
const Node *find(const Table &t, const std::string &path);

}  // namespace chtoml
