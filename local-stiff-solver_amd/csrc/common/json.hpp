// Minimal JSON reader for solver configuration documents (objects, arrays,
// numbers, strings, booleans, null).  The reference validates its documents with
// jse against linear-solver-spec.json (src/polysolve/linear/Solver.cpp:143-173);
// this reader only parses, Config applies the typed rules.
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lss::json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool is_int = false;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;  // keeps document order

    const Value *find(const std::string &k) const
    {
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class Parser {
public:
    explicit Parser(const std::string &s) : s_(s) {}
    Value parse()
    {
        Value v = value();
        ws();
        if (i_ != s_.size()) err("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t i_ = 0;

    [[noreturn]] void err(const char *m) { throw std::runtime_error(std::string("JSON: ") + m + " at offset " + std::to_string(i_)); }
    void ws()
    {
        while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
    }
    bool lit(const char *w)
    {
        size_t k = 0;
        while (w[k] && i_ + k < s_.size() && s_[i_ + k] == w[k]) ++k;
        if (w[k]) return false;
        i_ += k;
        return true;
    }
    Value value()
    {
        ws();
        if (i_ >= s_.size()) err("unexpected end");
        Value v;
        char c = s_[i_];
        if (c == '{') {
            v.kind = Value::Object;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return v;
            }
            while (true) {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') err("expected key");
                std::string k = string();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') err("expected ':'");
                ++i_;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    break;
                }
                err("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Value::Array;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    break;
                }
                err("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
        } else if (lit("null")) {
            v.kind = Value::Null;
        } else {
            const char *b = s_.c_str() + i_;
            char *e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) err("bad value");
            v.kind = Value::Number;
            v.is_int = true;
            for (const char *q = b; q < e; ++q)
                if (*q == '.' || *q == 'e' || *q == 'E') v.is_int = false;
            i_ += static_cast<size_t>(e - b);
        }
        return v;
    }
    std::string string()
    {
        ++i_;  // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) err("bad escape");
                char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': i_ += 4; out += '?'; break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) err("unterminated string");
        ++i_;
        return out;
    }
};

inline Value parse(const std::string &s) { return Parser(s).parse(); }

}  // namespace lss::json
