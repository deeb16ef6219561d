// toml_lite.cpp -- recursive-descent parser for the TOML subset of toml_lite.h.
#include "toml_lite.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace chtoml {

namespace {

struct Cursor {
    const std::string &t;
    size_t p = 0;
    int line = 1;
    std::string err;

    explicit Cursor(const std::string &text) : t(text) {}
    bool eof() const { return p >= t.size(); }
    char peek() const { return eof() ? '\0' : t[p]; }
    void skip_ws() {
        while(!eof() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\r')) p++;
    }
    void skip_comment() {
        if(peek() == '#')
            while(!eof() && t[p] != '\n') p++;
    }
    bool fail(const std::string &m) {
        if(err.empty()) err = "line " + std::to_string(line) + ": " + m;
        return false;
    }
};

bool parse_value(Cursor &c, Node &out);

bool parse_key_part(Cursor &c, std::string &key) {
    c.skip_ws();
    key.clear();
    if(c.peek() == '"' || c.peek() == '\'') {
        char q = c.t[c.p++];
        while(!c.eof() && c.peek() != q && c.peek() != '\n') key += c.t[c.p++];
        if(c.peek() != q) return c.fail("unterminated quoted key");
        c.p++;
        return true;
    }
    while(!c.eof() && (std::isalnum((unsigned char)c.peek()) || c.peek() == '_' || c.peek() == '-')) key += c.t[c.p++];
    if(key.empty()) return c.fail("expected a key");
    return true;
}

bool parse_key(Cursor &c, std::vector<std::string> &parts) {
    parts.clear();
    std::string k;
    if(!parse_key_part(c, k)) return false;
    parts.push_back(k);
    c.skip_ws();
    while(c.peek() == '.') {
        c.p++;
        if(!parse_key_part(c, k)) return false;
        parts.push_back(k);
        c.skip_ws();
    }
    return true;
}

bool parse_string(Cursor &c, Node &out) {
    char q = c.t[c.p++];
    std::string s;
    while(!c.eof() && c.peek() != q) {
        char ch = c.t[c.p++];
        if(ch == '\n') return c.fail("newline in string");
        if(q == '"' && ch == '\\') {
            char e = c.t[c.p++];
            switch(e) {
            case 'n': s += '\n'; break;
            case 't': s += '\t'; break;
            case '\\': s += '\\'; break;
            case '"': s += '"'; break;
            default: return c.fail("unsupported escape");
            }
        }
        else s += ch;
    }
    if(c.eof()) return c.fail("unterminated string");
    c.p++;
    out.kind = Node::STRING;
    out.s = s;
    return true;
}

bool parse_number_or_bool(Cursor &c, Node &out) {
    std::string tok;
    while(!c.eof()) {
        char ch = c.peek();
        if(std::isalnum((unsigned char)ch) || ch == '+' || ch == '-' || ch == '.' || ch == '_') tok += c.t[c.p++];
        else break;
    }
    if(tok == "true" || tok == "false") {
        out.kind = Node::BOOL;
        out.b = tok == "true";
        return true;
    }
    std::string clean;
    for(char ch : tok)
        if(ch != '_') clean += ch;
    if(clean == "inf" || clean == "+inf" || clean == "-inf" || clean == "nan" || clean == "+nan" || clean == "-nan") {
        out.kind = Node::FLOAT;
        out.f = clean.find("nan") != std::string::npos ? NAN : (clean[0] == '-' ? -INFINITY : INFINITY);
        return true;
    }
    if(clean.empty()) return c.fail("expected a value");
    bool is_float = clean.find_first_of(".eE") != std::string::npos;
    char *end = nullptr;
    if(is_float) {
        out.kind = Node::FLOAT;
        out.f = std::strtod(clean.c_str(), &end);
    }
    else {
        out.kind = Node::INT;
        out.i = std::strtoll(clean.c_str(), &end, 10);
    }
    if(!end || *end != '\0') return c.fail("bad value '" + tok + "'");
    return true;
}

bool insert(Cursor &c, Table &t, const std::vector<std::string> &parts, Node v) {
    Table *cur = &t;
    for(size_t i = 0; i + 1 < parts.size(); i++) {
        Node &n = (*cur)[parts[i]];
        if(n.kind == Node::NONE) {
            n.kind = Node::TABLE;
            n.tab = std::make_shared<Table>();
        }
        if(n.kind != Node::TABLE) return c.fail("key '" + parts[i] + "' is not a table");
        cur = n.tab.get();
    }
    if(cur->count(parts.back()) && (*cur)[parts.back()].kind != Node::TABLE)
        return c.fail("duplicate key '" + parts.back() + "'");
    (*cur)[parts.back()] = std::move(v);
    return true;
}

bool parse_inline_table(Cursor &c, Node &out) {
    c.p++;  // {
    out.kind = Node::TABLE;
    out.tab = std::make_shared<Table>();
    c.skip_ws();
    if(c.peek() == '}') {
        c.p++;
        return true;
    }
    while(true) {
        std::vector<std::string> parts;
        if(!parse_key(c, parts)) return false;
        c.skip_ws();
        if(c.peek() != '=') return c.fail("expected '=' in inline table");
        c.p++;
        c.skip_ws();
        Node v;
        if(!parse_value(c, v)) return false;
        if(!insert(c, *out.tab, parts, std::move(v))) return false;
        c.skip_ws();
        if(c.peek() == ',') {
            c.p++;
            continue;
        }
        if(c.peek() == '}') {
            c.p++;
            return true;
        }
        return c.fail("expected ',' or '}'");
    }
}

bool parse_array(Cursor &c, Node &out) {
    c.p++;  // [
    out.kind = Node::ARRAY;
    while(true) {
        while(!c.eof() && (std::isspace((unsigned char)c.peek()) || c.peek() == '#')) {
            if(c.peek() == '#') c.skip_comment();
            else {
                if(c.peek() == '\n') c.line++;
                c.p++;
            }
        }
        if(c.peek() == ']') {
            c.p++;
            return true;
        }
        Node v;
        if(!parse_value(c, v)) return false;
        out.arr.push_back(std::move(v));
        while(!c.eof() && std::isspace((unsigned char)c.peek())) {
            if(c.peek() == '\n') c.line++;
            c.p++;
        }
        if(c.peek() == ',') {
            c.p++;
            continue;
        }
        if(c.peek() == ']') {
            c.p++;
            return true;
        }
        return c.fail("expected ',' or ']'");
    }
}

bool parse_value(Cursor &c, Node &out) {
    c.skip_ws();
    char ch = c.peek();
    if(ch == '"' || ch == '\'') return parse_string(c, out);
    if(ch == '[') return parse_array(c, out);
    if(ch == '{') return parse_inline_table(c, out);
    return parse_number_or_bool(c, out);
}

}  // namespace

bool parse(const std::string &text, Table &out, std::string &err) {
    Cursor c(text);
    Table *current = &out;
    while(!c.eof()) {
        c.skip_ws();
        c.skip_comment();
        if(c.peek() == '\n') {
            c.p++;
            c.line++;
            continue;
        }
        if(c.eof()) break;
        if(c.peek() == '[') {
            c.p++;
            // [[a.b]]: append a new table to the array a.b and fill that one
            const bool array_of_tables = c.peek() == '[';
            if(array_of_tables) c.p++;
            std::vector<std::string> parts;
            if(!parse_key(c, parts)) break;
            if(c.peek() != ']' || (array_of_tables && (c.p + 1 >= c.t.size() || c.t[c.p + 1] != ']'))) {
                c.fail(array_of_tables ? "expected ']]'" : "expected ']'");
                break;
            }
            c.p += array_of_tables ? 2 : 1;
            current = &out;
            bool ok = true;
            for(size_t i = 0; i < parts.size(); i++) {
                Node &n = (*current)[parts[i]];
                const bool last = i + 1 == parts.size();
                if(array_of_tables && last) {
                    if(n.kind == Node::NONE) n.kind = Node::ARRAY;
                    if(n.kind != Node::ARRAY) {
                        ok = c.fail("'" + parts[i] + "' is not an array of tables");
                        break;
                    }
                    Node t;
                    t.kind = Node::TABLE;
                    t.tab = std::make_shared<Table>();
                    n.arr.push_back(t);
                    current = n.arr.back().tab.get();
                    break;
                }
                if(n.kind == Node::NONE) {
                    n.kind = Node::TABLE;
                    n.tab = std::make_shared<Table>();
                }
                // a dotted header through an array of tables continues in its last element
                if(n.kind == Node::ARRAY && !n.arr.empty() && n.arr.back().kind == Node::TABLE) {
                    current = n.arr.back().tab.get();
                    continue;
                }
                if(n.kind != Node::TABLE) {
                    ok = c.fail("'" + parts[i] + "' is not a table");
                    break;
                }
                current = n.tab.get();
            }
            if(!ok) break;
        }
        else {
            std::vector<std::string> parts;
            if(!parse_key(c, parts)) break;
            c.skip_ws();
            if(c.peek() != '=') {
                c.fail("expected '='");
                break;
            }
            c.p++;
            Node v;
            if(!parse_value(c, v)) break;
            if(!insert(c, *current, parts, std::move(v))) break;
        }
        c.skip_ws();
        c.skip_comment();
        if(!c.eof() && c.peek() != '\n') {
            c.fail("unexpected text after value");
            break;
        }
    }
    err = c.err;
    return err.empty();
}

bool parse_file(const std::string &path, Table &out, std::string &err) {
    std::ifstream in(path);
    if(!in) {
        err = "cannot open '" + path + "'";
        return false;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str(), out, err);
}

const Node *find(const Table &t, const std::string &path) {
    const Table *cur = &t;
    size_t start = 0;
    while(true) {
        size_t dot = path.find('.', start);
        std::string k = path.substr(start, dot == std::string::npos ? std::string::npos : dot - start);
        auto it = cur->find(k);
        if(it == cur->end()) return nullptr;
        if(dot == std::string::npos) return &it->second;
        if(it->second.kind != Node::TABLE) return nullptr;
        cur = it->second.tab.get();
        start = dot + 1;
    }
}

}  // namespace chtoml
