// json_lite.h -- minimal JSON reader for Lucid scene files.
//
// A DOM reader with one twist for large meshes: an array whose leaves are all
// numbers (e.g. "vertices": [[x,y,z], ...]) is stored flattened as one
// std::vector<double> plus its shape, instead of one node per number.
#pragma once

#include <charconv>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace lucid {
namespace json {

struct Value {
    enum Kind : uint8_t { Null, Bool, Num, Str, Arr, Obj, NumArr };
    Kind kind = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Value> arr;                              // Arr
    std::vector<std::pair<std::string, Value>> obj;      // Obj
    std::vector<double> flat;                            // NumArr leaves
    std::vector<int64_t> shape;                          // NumArr shape (regular)
    bool regular = true;

    const Value* get(const char* key) const {
        if (kind != Obj) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool has(const char* key) const { return get(key) != nullptr; }
    bool is_array() const { return kind == Arr || kind == NumArr; }
    bool is_object() const { return kind == Obj; }
    bool is_null() const { return kind == Null; }
    int64_t size() const {
        if (kind == Arr) return (int64_t)arr.size();
        if (kind == NumArr) return shape.empty() ? 0 : shape[0];
        return 0;
    }
};

class Parser {
public:
    Parser(const char* p, size_t n) : p_(p), e_(p + n) {}

    Value parse() {
        Value v = value();
        ws();
        if (p_ != e_) fail("trailing characters");
        return v;
    }

private:
    const char* p_;
    const char* e_;

    [[noreturn]] void fail(const char* what) { throw std::runtime_error(std::string("JSON parse error: ") + what); }
    void ws() {
        while (p_ < e_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
    }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if ((size_t)(e_ - p_) >= n && std::memcmp(p_, w, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    double number() {
        const char* s = p_;
        if (p_ < e_ && (*p_ == '-' || *p_ == '+')) ++p_;
        while (p_ < e_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '-' ||
                           *p_ == '+'))
            ++p_;
        double d = 0.0;
        const char* b = (*s == '+') ? s + 1 : s;
        auto r = std::from_chars(b, p_, d);
        if (r.ec != std::errc() || r.ptr != p_) fail("bad number");
        return d;
    }
    std::string str() {
        ++p_;  // opening quote
        std::string out;
        while (p_ < e_ && *p_ != '"') {
            char c = *p_++;
            if (c == '\\') {
                if (p_ >= e_) fail("bad escape");
                char x = *p_++;
                switch (x) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        if (e_ - p_ < 4) fail("bad \\u escape");
                        unsigned cp = 0;
                        for (int i = 0; i < 4; ++i) {
                            char h = *p_++;
                            cp <<= 4;
                            if (h >= '0' && h <= '9') cp |= h - '0';
                            else if (h >= 'a' && h <= 'f') cp |= h - 'a' + 10;
                            else if (h >= 'A' && h <= 'F') cp |= h - 'A' + 10;
                            else fail("bad \\u escape");
                        }
                        if (cp >= 0xD800 && cp <= 0xDBFF && e_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                            unsigned lo = 0;
                            const char* q = p_ + 2;
                            for (int i = 0; i < 4; ++i) {
                                char h = q[i];
                                lo <<= 4;
                                if (h >= '0' && h <= '9') lo |= h - '0';
                                else if (h >= 'a' && h <= 'f') lo |= h - 'a' + 10;
                                else if (h >= 'A' && h <= 'F') lo |= h - 'A' + 10;
                            }
                            if (lo >= 0xDC00 && lo <= 0xDFFF) {
                                cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                                p_ += 6;
                            }
                        }
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                        else if (cp < 0x10000) {
                            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (p_ >= e_) fail("unterminated string");
        ++p_;
        return out;
    }
    // Numeric-only array at p_ (which points at '['): appends leaves to flat,
    // records the shape of this level; returns false (position undefined)
    // when a non-number leaf appears.
    bool num_array(std::vector<double>& flat, std::vector<int64_t>& shape, size_t level, bool& regular) {
        ++p_;
        ws();
        int64_t count = 0;
        if (p_ < e_ && *p_ == ']') {
            ++p_;
        } else {
            for (;;) {
                ws();
                if (p_ >= e_) return false;
                if (*p_ == '[') {
                    if (!num_array(flat, shape, level + 1, regular)) return false;
                } else if (*p_ == '-' || (*p_ >= '0' && *p_ <= '9')) {
                    if (shape.size() > level + 1) regular = false;  // mixed depth
                    flat.push_back(number());
                } else {
                    return false;
                }
                ++count;
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == ']') { ++p_; break; }
                return false;
            }
        }
        if (shape.size() <= level) shape.resize(level + 1, -1);
        if (shape[level] < 0) shape[level] = count;
        else if (level > 0 && shape[level] != count) regular = false;
        if (level == 0) shape[0] = count;
        return true;
    }
    Value value() {
        ws();
        if (p_ >= e_) fail("unexpected end");
        Value v;
        char c = *p_;
        if (c == '{') {
            v.kind = Value::Obj;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (p_ >= e_ || *p_ != '"') fail("expected key");
                std::string k = str();
                ws();
                if (p_ >= e_ || *p_ != ':') fail("expected ':'");
                ++p_;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == '}') { ++p_; break; }
                fail("expected ',' or '}'");
            }
            return v;
        }
        if (c == '[') {
            const char* save = p_;
            Value nv;
            nv.kind = Value::NumArr;
            if (num_array(nv.flat, nv.shape, 0, nv.regular)) return nv;
            p_ = save;
            v.kind = Value::Arr;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == ']') { ++p_; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == ']') { ++p_; break; }
                fail("expected ',' or ']'");
            }
            return v;
        }
        if (c == '"') { v.kind = Value::Str; v.s = str(); return v; }
        if (lit("true")) { v.kind = Value::Bool; v.b = true; return v; }
        if (lit("false")) { v.kind = Value::Bool; v.b = false; return v; }
        if (lit("null")) { v.kind = Value::Null; return v; }
        if (c == '-' || (c >= '0' && c <= '9')) { v.kind = Value::Num; v.n = number(); return v; }
        fail("unexpected character");
    }
};

inline Value parse(const std::string& text) { return Parser(text.data(), text.size()).parse(); }

// Accessors with nlohmann-like strictness: wrong types throw.
inline double num(const Value& v) {
    if (v.kind == Value::Num) return v.n;
    if (v.kind == Value::Bool) return v.b ? 1.0 : 0.0;
    throw std::runtime_error("JSON: expected a number");
}
inline bool boolean(const Value& v) {
    if (v.kind == Value::Bool) return v.b;
    throw std::runtime_error("JSON: expected a boolean");
}
inline const std::string& string(const Value& v) {
    if (v.kind == Value::Str) return v.s;
    throw std::runtime_error("JSON: expected a string");
}
// i-th number of a (flattened) numeric array of at least i+1 leaves
inline double at(const Value& v, size_t i) {
    if (v.kind == Value::NumArr && i < v.flat.size()) return v.flat[i];
    if (v.kind == Value::Arr && i < v.arr.size()) return num(v.arr[i]);
    throw std::runtime_error("JSON: array too short or not numeric");
}

}  // namespace json
}  // namespace lucid
