// JSON value, parser and writer for the research wire format (host code).
//
// The reference reads and writes its requests, outputs and checkpoints with
// serde_json (features "float_roundtrip", no "preserve_order"; core/Cargo.toml):
//   * objects are maps with sorted keys (a later duplicate key replaces an
//     earlier one) and are written in key order;
//   * an integer literal is kept as an unsigned or signed 64-bit integer when
//     it fits, any other number is an f64 parsed with correct rounding;
//   * f64 values are written with the shortest digits that round-trip, in the
//     layout of the ryu crate (1.0, 0.001, 1e-7, 1.25e30), and non-finite
//     values as null;
//   * nesting deeper than 128 levels is an error.
// Numbers therefore survive a write/parse cycle bit for bit.
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <system_error>
#include <utility>
#include <vector>

namespace mpm_json {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Value {
    enum Kind { Null, Bool, UInt, Int, Float, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    uint64_t u = 0;
    int64_t i = 0;
    double d = 0.0;
    std::string s;
    std::vector<Value> a;
    std::map<std::string, Value> o;

    Value() = default;
    static Value null() { return Value(); }
    static Value boolean(bool v) {
        Value r;
        r.kind = Bool;
        r.b = v;
        return r;
    }
    static Value uint(uint64_t v) {
        Value r;
        r.kind = UInt;
        r.u = v;
        return r;
    }
    static Value integer(int64_t v) {
        if (v >= 0) return uint((uint64_t)v);
        Value r;
        r.kind = Int;
        r.i = v;
        return r;
    }
    // An f64; a non-finite value is written as null, as serde_json does.
    static Value number(double v) {
        Value r;
        r.kind = Float;
        r.d = v;
        return r;
    }
    static Value string(std::string v) {
        Value r;
        r.kind = String;
        r.s = std::move(v);
        return r;
    }
    static Value array() {
        Value r;
        r.kind = Array;
        return r;
    }
    static Value object() {
        Value r;
        r.kind = Object;
        return r;
    }
    static Value vec(const double* v, size_t n) {
        Value r = array();
        r.a.reserve(n);
        for (size_t k = 0; k < n; ++k) r.a.push_back(number(v[k]));
        return r;
    }

    bool is_number() const { return kind == UInt || kind == Int || kind == Float; }
    const Value* find(const std::string& key) const {
        if (kind != Object) return nullptr;
        auto it = o.find(key);
        return it == o.end() ? nullptr : &it->second;
    }
    Value& operator[](const std::string& key) {
        if (kind != Object) throw Error("not an object");
        return o[key];
    }
    // serde_json::Value equality: numbers compare by representation class
    // (integer 3 differs from float 3.0), floats by value.
    bool operator==(const Value& r) const {
        if (kind != r.kind) return false;
        switch (kind) {
            case Null: return true;
            case Bool: return b == r.b;
            case UInt: return u == r.u;
            case Int: return i == r.i;
            case Float: return d == r.d;
            case String: return s == r.s;
            case Array: return a == r.a;
            case Object: return o == r.o;
        }
        return false;
    }
    bool operator!=(const Value& r) const { return !(*this == r); }
};

// ------------------------------------------------------------------ writer

// f64 in the ryu crate's layout with the shortest round-trip digits.
inline void write_f64(std::string& out, double v) {
    if (!std::isfinite(v)) {
        out += "null";
        return;
    }
    if (v == 0.0) {
        out += std::signbit(v) ? "-0.0" : "0.0";
        return;
    }
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    if (res.ec != std::errc()) throw Error("float formatting failed");
    // buf = [-]d[.ddd]e(+|-)XX
    std::string t(buf, res.ptr);
    size_t pos = 0;
    if (t[0] == '-') {
        out += '-';
        pos = 1;
    }
    size_t e = t.find('e', pos);
    std::string digits;
    for (size_t k = pos; k < e; ++k)
        if (t[k] != '.') digits += t[k];
    int exp10 = std::stoi(t.substr(e + 1));
    const int length = (int)digits.size();
    const int k = exp10 - (length - 1);  // value = digits x 10^k
    const int kk = length + k;           // 10^(kk-1) <= |v| < 10^kk
    if (0 <= k && kk <= 16) {
        out += digits;
        out.append((size_t)(kk - length), '0');
        out += ".0";
    } else if (0 < kk && kk <= 16) {
        out += digits.substr(0, (size_t)kk);
        out += '.';
        out += digits.substr((size_t)kk);
    } else if (-5 < kk && kk <= 0) {
        out += "0.";
        out.append((size_t)(-kk), '0');
        out += digits;
    } else if (length == 1) {
        out += digits;
        out += 'e';
        out += std::to_string(kk - 1);
    } else {
        out += digits[0];
        out += '.';
        out += digits.substr(1);
        out += 'e';
        out += std::to_string(kk - 1);
    }
}

inline void write_string(std::string& out, const std::string& s) {
    static const char* hex = "0123456789abcdef";
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            default:
                if (c < 0x20) {
                    out += "\\u00";
                    out += hex[c >> 4];
                    out += hex[c & 15];
                } else {
                    out += (char)c;
                }
        }
    }
    out += '"';
}

inline void write(std::string& out, const Value& v) {
    switch (v.kind) {
        case Value::Null: out += "null"; break;
        case Value::Bool: out += v.b ? "true" : "false"; break;
        case Value::UInt: out += std::to_string(v.u); break;
        case Value::Int: out += std::to_string(v.i); break;
        case Value::Float: write_f64(out, v.d); break;
        case Value::String: write_string(out, v.s); break;
        case Value::Array: {
            out += '[';
            for (size_t k = 0; k < v.a.size(); ++k) {
                if (k) out += ',';
                write(out, v.a[k]);
            }
            out += ']';
            break;
        }
        case Value::Object: {
            out += '{';
            bool first = true;
            for (const auto& kv : v.o) {
                if (!first) out += ',';
                first = false;
                write_string(out, kv.first);
                out += ':';
                write(out, kv.second);
            }
            out += '}';
            break;
        }
    }
}

inline std::string dump(const Value& v) {
    std::string out;
    write(out, v);
    return out;
}

// ------------------------------------------------------------------ parser

class Parser {
public:
    Parser(const char* text, size_t n) : p_(text), end_(text + n) {}

    Value parse_document() {
        Value v = parse_value(0);
        skip_ws();
        if (p_ != end_) fail("trailing characters");
        return v;
    }

private:
    const char* p_;
    const char* end_;
    const char* begin_ = p_;

    [[noreturn]] void fail(const char* what) const {
        throw Error(std::string(what) + " at offset " + std::to_string((size_t)(p_ - begin_)));
    }
    void skip_ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
    }
    bool literal(const char* word) {
        size_t n = std::strlen(word);
        if ((size_t)(end_ - p_) >= n && std::memcmp(p_, word, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    Value parse_value(int depth) {
        if (depth > 128) fail("recursion limit exceeded");
        skip_ws();
        if (p_ == end_) fail("EOF while parsing a value");
        char c = *p_;
        if (c == '{') return parse_object(depth);
        if (c == '[') return parse_array(depth);
        if (c == '"') return Value::string(parse_string());
        if (c == '-' || (c >= '0' && c <= '9')) return parse_number();
        if (literal("null")) return Value::null();
        if (literal("true")) return Value::boolean(true);
        if (literal("false")) return Value::boolean(false);
        fail("expected value");
    }

    Value parse_object(int depth) {
        ++p_;
        Value v = Value::object();
        skip_ws();
        if (p_ < end_ && *p_ == '}') {
            ++p_;
            return v;
        }
        for (;;) {
            skip_ws();
            if (p_ == end_ || *p_ != '"') fail("key must be a string");
            std::string key = parse_string();
            skip_ws();
            if (p_ == end_ || *p_ != ':') fail("expected ':'");
            ++p_;
            v.o[key] = parse_value(depth + 1);
            skip_ws();
            if (p_ == end_) fail("EOF while parsing an object");
            if (*p_ == ',') {
                ++p_;
                continue;
            }
            if (*p_ == '}') {
                ++p_;
                return v;
            }
            fail("expected ',' or '}'");
        }
    }

    Value parse_array(int depth) {
        ++p_;
        Value v = Value::array();
        skip_ws();
        if (p_ < end_ && *p_ == ']') {
            ++p_;
            return v;
        }
        for (;;) {
            v.a.push_back(parse_value(depth + 1));
            skip_ws();
            if (p_ == end_) fail("EOF while parsing a list");
            if (*p_ == ',') {
                ++p_;
                continue;
            }
            if (*p_ == ']') {
                ++p_;
                return v;
            }
            fail("expected ',' or ']'");
        }
    }

    static void put_utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (end_ - p_ < 4) fail("EOF while parsing a string");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9')
                v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f')
                v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F')
                v |= (uint32_t)(c - 'A' + 10);
            else
                fail("invalid escape");
        }
        return v;
    }

    std::string parse_string() {
        ++p_;
        std::string s;
        for (;;) {
            if (p_ == end_) fail("EOF while parsing a string");
            unsigned char c = (unsigned char)*p_++;
            if (c == '"') return s;
            if (c < 0x20) fail("control character while parsing a string");
            if (c != '\\') {
                s += (char)c;
                continue;
            }
            if (p_ == end_) fail("EOF while parsing a string");
            char e = *p_++;
            switch (e) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        if (end_ - p_ < 2 || p_[0] != '\\' || p_[1] != 'u') fail("lone leading surrogate");
                        p_ += 2;
                        uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("invalid surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("lone trailing surrogate");
                    }
                    put_utf8(s, cp);
                    break;
                }
                default: fail("invalid escape");
            }
        }
    }

    Value parse_number() {
        const char* start = p_;
        bool neg = false;
        if (*p_ == '-') {
            neg = true;
            ++p_;
        }
        if (p_ == end_ || !(*p_ >= '0' && *p_ <= '9')) fail("invalid number");
        if (*p_ == '0') {
            ++p_;
            if (p_ < end_ && *p_ >= '0' && *p_ <= '9') fail("invalid number");
        } else {
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        bool is_float = false;
        if (p_ < end_ && *p_ == '.') {
            is_float = true;
            ++p_;
            if (p_ == end_ || !(*p_ >= '0' && *p_ <= '9')) fail("invalid number");
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (p_ < end_ && (*p_ == 'e' || *p_ == 'E')) {
            is_float = true;
            ++p_;
            if (p_ < end_ && (*p_ == '+' || *p_ == '-')) ++p_;
            if (p_ == end_ || !(*p_ >= '0' && *p_ <= '9')) fail("invalid number");
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (!is_float) {
            if (neg) {
                int64_t v;
                auto r = std::from_chars(start, p_, v);
                if (r.ec == std::errc() && r.ptr == p_) return Value::integer(v);
            } else {
                uint64_t v;
                auto r = std::from_chars(start, p_, v);
                if (r.ec == std::errc() && r.ptr == p_) return Value::uint(v);
            }
        }
        double d = 0.0;
        auto r = std::from_chars(start, p_, d);
        if (r.ptr != p_) fail("invalid number");
        if (r.ec == std::errc::result_out_of_range) {
            // Subnormal results and underflow to zero are values, as in
            // serde_json; overflow is an error.  strtod rounds both correctly.
            const std::string text(start, p_);
            d = std::strtod(text.c_str(), nullptr);
        } else if (r.ec != std::errc()) {
            fail("invalid number");
        }
        if (!std::isfinite(d)) fail("number out of range");
        return Value::number(d);
    }
};

inline Value parse(const std::string& text) {
    Parser p(text.data(), text.size());
    return p.parse_document();
}

}  // namespace mpm_json
