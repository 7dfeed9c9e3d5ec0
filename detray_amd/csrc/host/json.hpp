// Minimal JSON document model for the detector files: a recursive-descent
// parser (RFC 8259: objects keep their key order, integers are kept exact
// as 64-bit values next to their double value, strings are UTF-8 with
// \uXXXX escapes decoded) and a pretty printer with 4-space indentation,
// one array element per line, doubles in the shortest form that reads back
// to the same value.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dtr_json {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Value {
    enum Kind { Null, Bool, Int, UInt, Float, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    int64_t i = 0;   // Int: negative integers
    uint64_t u = 0;  // UInt: non-negative integers
    double d = 0.0;  // every number
    std::string s;
    std::vector<Value> a;
    std::vector<std::pair<std::string, Value>> o;

    static Value null() { return Value{}; }
    static Value boolean(bool v) {
        Value x;
        x.kind = Bool;
        x.b = v;
        return x;
    }
    static Value uint(uint64_t v) {
        Value x;
        x.kind = UInt;
        x.u = v;
        x.d = static_cast<double>(v);
        return x;
    }
    static Value number(double v) {
        Value x;
        x.kind = Float;
        x.d = v;
        return x;
    }
    static Value string(std::string v) {
        Value x;
        x.kind = String;
        x.s = std::move(v);
        return x;
    }
    static Value array() {
        Value x;
        x.kind = Array;
        return x;
    }
    static Value object() {
        Value x;
        x.kind = Object;
        return x;
    }

    bool is_number() const { return kind == Int || kind == UInt || kind == Float; }

    const Value* find(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    Value& set(const char* key, Value v) {
        o.emplace_back(key, std::move(v));
        return o.back().second;
    }
    Value& push(Value v) {
        a.push_back(std::move(v));
        return a.back();
    }
};

// ---------------------------------------------------------------------------
// Parser
// ---------------------------------------------------------------------------
class Parser {
  public:
    Parser(const char* begin, const char* end) : p_(begin), b_(begin), e_(end) {}

    Value document() {
        Value v = value(0);
        ws();
        if (p_ != e_) fail("trailing characters after the document");
        return v;
    }

  private:
    const char* p_;
    const char* b_;
    const char* e_;

    [[noreturn]] void fail(const char* what) const {
        size_t line = 1, col = 1;
        for (const char* q = b_; q < p_ && q < e_; ++q) {
            if (*q == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw Error("JSON parse error at line " + std::to_string(line) + ", column " +
                    std::to_string(col) + ": " + what);
    }

    void ws() {
        while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }

    bool literal(const char* word) {
        const size_t n = std::strlen(word);
        if (static_cast<size_t>(e_ - p_) >= n && std::memcmp(p_, word, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    Value value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (p_ == e_) fail("unexpected end of input");
        const char c = *p_;
        if (c == '{') return object(depth);
        if (c == '[') return array(depth);
        if (c == '"') return Value::string(string());
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        if (literal("true")) return Value::boolean(true);
        if (literal("false")) return Value::boolean(false);
        if (literal("null")) return Value::null();
        fail("unexpected character");
    }

    Value object(int depth) {
        Value v = Value::object();
        ++p_;
        ws();
        if (p_ < e_ && *p_ == '}') {
            ++p_;
            return v;
        }
        for (;;) {
            ws();
            if (p_ == e_ || *p_ != '"') fail("expected a key string");
            std::string key = string();
            ws();
            if (p_ == e_ || *p_ != ':') fail("expected ':'");
            ++p_;
            v.o.emplace_back(std::move(key), value(depth + 1));
            ws();
            if (p_ == e_) fail("unexpected end of input in an object");
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

    Value array(int depth) {
        Value v = Value::array();
        ++p_;
        ws();
        if (p_ < e_ && *p_ == ']') {
            ++p_;
            return v;
        }
        for (;;) {
            v.a.push_back(value(depth + 1));
            ws();
            if (p_ == e_) fail("unexpected end of input in an array");
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

    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xc0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3f));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xe0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3f));
            out += static_cast<char>(0x80 | (cp & 0x3f));
        } else {
            out += static_cast<char>(0xf0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3f));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3f));
            out += static_cast<char>(0x80 | (cp & 0x3f));
        }
    }

    uint32_t hex4() {
        if (e_ - p_ < 4) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<uint32_t>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<uint32_t>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<uint32_t>(c - 'A' + 10);
            else fail("bad hex digit in \\u escape");
        }
        return v;
    }

    std::string string() {
        ++p_;  // opening quote
        std::string out;
        for (;;) {
            if (p_ == e_) fail("unterminated string");
            const char c = *p_++;
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ == e_) fail("unterminated escape");
            const char esc = *p_++;
            switch (esc) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xd800 && cp < 0xdc00) {
                        if (e_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u') fail("unpaired surrogate");
                        p_ += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xdc00 || lo >= 0xe000) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                    } else if (cp >= 0xdc00 && cp < 0xe000) {
                        fail("unpaired surrogate");
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }

    Value number() {
        const char* start = p_;
        bool neg = false, integral = true;
        if (*p_ == '-') {
            neg = true;
            ++p_;
        }
        if (p_ == e_ || !(*p_ >= '0' && *p_ <= '9')) fail("bad number");
        if (*p_ == '0') {
            ++p_;
        } else {
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (p_ < e_ && *p_ == '.') {
            integral = false;
            ++p_;
            if (p_ == e_ || !(*p_ >= '0' && *p_ <= '9')) fail("bad fraction");
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
            integral = false;
            ++p_;
            if (p_ < e_ && (*p_ == '+' || *p_ == '-')) ++p_;
            if (p_ == e_ || !(*p_ >= '0' && *p_ <= '9')) fail("bad exponent");
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        const std::string text(start, p_);
        Value v;
        v.d = std::strtod(text.c_str(), nullptr);
        if (integral) {
            errno = 0;
            if (neg) {
                const long long x = std::strtoll(text.c_str(), nullptr, 10);
                if (errno == ERANGE) fail("integer out of range");
                v.kind = Value::Int;
                v.i = x;
                if (x == 0) {
                    v.kind = Value::UInt;
                    v.u = 0;
                }
            } else {
                const unsigned long long x = std::strtoull(text.c_str(), nullptr, 10);
                if (errno == ERANGE) fail("integer out of range");
                v.kind = Value::UInt;
                v.u = x;
            }
        } else {
            v.kind = Value::Float;
        }
        return v;
    }
};

inline Value parse(const std::string& text) {
    return Parser(text.data(), text.data() + text.size()).document();
}

// ---------------------------------------------------------------------------
// Printer
// ---------------------------------------------------------------------------
inline std::string format_double(double x) {
    if (!std::isfinite(x)) return "null";
    char buf[40];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*g", prec, x);
        if (std::strtod(buf, nullptr) == x) break;
    }
    std::string s(buf);
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    return s;
}

inline void escape(std::string& out, const std::string& s) {
    out += '"';
    for (const char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (static_cast<unsigned char>(c) < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof(buf), "\\u%04x", static_cast<unsigned>(c));
                    out += buf;
                } else {
                    out += c;
                }
        }
    }
    out += '"';
}

inline void dump(std::string& out, const Value& v, int indent) {
    const std::string pad(static_cast<size_t>(indent + 4), ' ');
    const std::string close(static_cast<size_t>(indent), ' ');
    switch (v.kind) {
        case Value::Null: out += "null"; break;
        case Value::Bool: out += v.b ? "true" : "false"; break;
        case Value::Int: out += std::to_string(v.i); break;
        case Value::UInt: out += std::to_string(v.u); break;
        case Value::Float: out += format_double(v.d); break;
        case Value::String: escape(out, v.s); break;
        case Value::Array:
            if (v.a.empty()) {
                out += "[]";
                break;
            }
            out += "[\n";
            for (size_t k = 0; k < v.a.size(); ++k) {
                out += pad;
                dump(out, v.a[k], indent + 4);
                out += (k + 1 < v.a.size()) ? ",\n" : "\n";
            }
            out += close + "]";
            break;
        case Value::Object:
            if (v.o.empty()) {
                out += "{}";
                break;
            }
            out += "{\n";
            for (size_t k = 0; k < v.o.size(); ++k) {
                out += pad;
                escape(out, v.o[k].first);
                out += ": ";
                dump(out, v.o[k].second, indent + 4);
                out += (k + 1 < v.o.size()) ? ",\n" : "\n";
            }
            out += close + "}";
            break;
    }
}

inline std::string dump(const Value& v) {
    std::string out;
    dump(out, v, 0);
    out += '\n';
    return out;
}

}  // namespace dtr_json
