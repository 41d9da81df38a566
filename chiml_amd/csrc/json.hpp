// Minimal JSON reader for chiML input decks.
//
// The reference reads its deck with boost::property_tree after deleting every
// "//" comment to the end of the line (src/INPUTS/parallelInputs.cpp:1625-1649,
// src/main.cpp:26-40).  This reader does the same comment stripping and then
// parses standard JSON into a small tree.  Lookups use dotted paths like the
// property tree ("CompCell.size"), and missing optional keys fall back to the
// caller's default exactly as ptree::get<T>(path, default) does.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace chiml {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;  // insertion order kept

    bool is_null() const { return kind == Null; }
    bool is_array() const { return kind == Array; }
    bool is_object() const { return kind == Object; }

    const Value* find(const std::string& key) const {
        if (kind != Object) return nullptr;
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    // Dotted-path lookup ("CompCell.size"); nullptr when any part is absent.
    const Value* path(const std::string& p) const {
        const Value* cur = this;
        size_t start = 0;
        while (cur) {
            size_t dot = p.find('.', start);
            std::string part = p.substr(start, dot == std::string::npos ? std::string::npos : dot - start);
            cur = cur->find(part);
            if (dot == std::string::npos) break;
            start = dot + 1;
        }
        return cur;
    }
    // Children of an array or object (property-tree style iteration).
    size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }
    const Value& child(size_t i) const { return kind == Array ? arr.at(i) : obj.at(i).second; }

    double as_num() const {
        if (kind == Number) return num;
        if (kind == Bool) return b ? 1.0 : 0.0;
        if (kind == String) {
            char* end = nullptr;
            double v = std::strtod(str.c_str(), &end);
            if (end && *end == '\0' && !str.empty()) return v;
        }
        throw std::logic_error("JSON value is not a number");
    }
    bool as_bool() const {
        if (kind == Bool) return b;
        if (kind == String) {
            if (str == "true") return true;
            if (str == "false") return false;
        }
        if (kind == Number) return num != 0.0;
        throw std::logic_error("JSON value is not a bool");
    }
    std::string as_str() const {
        if (kind == String) return str;
        if (kind == Number) {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%.17g", num);
            return buf;
        }
        if (kind == Bool) return b ? "true" : "false";
        throw std::logic_error("JSON value is not a string");
    }
};

class Parser {
  public:
    explicit Parser(const std::string& text) : s_(strip_comments(text)) {}
    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }
    // Delete everything after "//" on each line, as the reference's
    // stripComments does before the deck reaches the JSON parser.
    static std::string strip_comments(const std::string& in) {
        std::string out;
        out.reserve(in.size());
        size_t pos = 0;
        while (pos < in.size()) {
            size_t eol = in.find('\n', pos);
            std::string line = in.substr(pos, eol == std::string::npos ? std::string::npos : eol - pos);
            size_t f = line.find('/');
            if (f != std::string::npos && f + 1 < line.size() && line[f + 1] == '/') line.erase(f);
            out += line;
            out += '\n';
            if (eol == std::string::npos) break;
            pos = eol + 1;
        }
        return out;
    }

  private:
    std::string s_;
    size_t i_ = 0;
    [[noreturn]] void fail(const char* what) {
        throw std::logic_error(std::string("JSON parse error: ") + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
    }
    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        char c = s_[i_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            Value v;
            v.kind = Value::String;
            v.str = string();
            return v;
        }
        if (s_.compare(i_, 4, "true") == 0) { i_ += 4; Value v; v.kind = Value::Bool; v.b = true; return v; }
        if (s_.compare(i_, 5, "false") == 0) { i_ += 5; Value v; v.kind = Value::Bool; v.b = false; return v; }
        if (s_.compare(i_, 4, "null") == 0) { i_ += 4; return Value(); }
        return number();
    }
    Value number() {
        const char* start = s_.c_str() + i_;
        char* end = nullptr;
        double d = std::strtod(start, &end);
        if (end == start) fail("bad value");
        i_ += static_cast<size_t>(end - start);
        Value v;
        v.kind = Value::Number;
        v.num = d;
        return v;
    }
    std::string string() {
        ++i_;  // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) fail("bad unicode escape");
                        unsigned cp = std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) out += static_cast<char>(cp);
                        else out += '?';
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
    Value array() {
        Value v;
        v.kind = Value::Array;
        ++i_;
        ws();
        if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
        while (true) {
            v.arr.push_back(value());
            ws();
            if (i_ >= s_.size()) fail("unterminated array");
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == ']') { ++i_; break; }
            fail("expected , or ]");
        }
        return v;
    }
    Value object() {
        Value v;
        v.kind = Value::Object;
        ++i_;
        ws();
        if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
        while (true) {
            ws();
            if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
            std::string key = string();
            ws();
            if (i_ >= s_.size() || s_[i_] != ':') fail("expected :");
            ++i_;
            v.obj.emplace_back(key, value());
            ws();
            if (i_ >= s_.size()) fail("unterminated object");
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == '}') { ++i_; break; }
            fail("expected , or }");
        }
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

// ptree-style accessors ---------------------------------------------------------
inline double get_num(const Value& v, const std::string& p) {
    const Value* x = v.path(p);
    if (!x) throw std::logic_error("missing required input key '" + p + "'");
    return x->as_num();
}
inline double get_num(const Value& v, const std::string& p, double def) {
    const Value* x = v.path(p);
    return (x && !x->is_null()) ? x->as_num() : def;
}
inline int get_int(const Value& v, const std::string& p, int def) {
    const Value* x = v.path(p);
    return (x && !x->is_null()) ? static_cast<int>(x->as_num()) : def;
}
inline bool get_bool(const Value& v, const std::string& p, bool def) {
    const Value* x = v.path(p);
    return (x && !x->is_null()) ? x->as_bool() : def;
}
inline std::string get_str(const Value& v, const std::string& p) {
    const Value* x = v.path(p);
    if (!x) throw std::logic_error("missing required input key '" + p + "'");
    return x->as_str();
}
inline std::string get_str(const Value& v, const std::string& p, const std::string& def) {
    const Value* x = v.path(p);
    return (x && !x->is_null()) ? x->as_str() : def;
}
// Child list (array or object); an absent key is an empty list.
inline const Value& get_child_or_empty(const Value& v, const std::string& p) {
    static const Value empty;
    const Value* x = v.path(p);
    return x ? *x : empty;
}
// as_ptArr: first three children as doubles, default triple when absent
// (src/INPUTS/parallelInputs.hpp:382-400).
inline void get_arr3(const Value& v, const std::string& p, double out[3], double def = 0.0) {
    const Value* x = v.path(p);
    if (!x || x->size() == 0) {
        out[0] = out[1] = out[2] = def;
        return;
    }
    out[0] = out[1] = out[2] = 0.0;
    for (size_t i = 0; i < x->size() && i < 3; ++i) out[i] = x->child(i).as_num();
}
inline std::vector<double> get_vec(const Value& v, const std::string& p) {
    std::vector<double> r;
    const Value* x = v.path(p);
    if (!x) return r;
    for (size_t i = 0; i < x->size(); ++i) r.push_back(x->child(i).as_num());
    return r;
}

}  // namespace json
}  // namespace chiml
