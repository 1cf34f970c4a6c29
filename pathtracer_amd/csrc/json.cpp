// json.cpp — see json.hpp.
#include "json.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace ptjson {

const Value* Value::get(const std::string& key) const {
    if (kind != Kind::Object) return nullptr;
    for (const auto& kv : obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

size_t Value::size() const {
    if (kind == Kind::Array) return arr.size();
    if (kind == Kind::Object) return obj.size();
    return 0;
}

const Value& Value::at(size_t idx) const {
    if (kind != Kind::Array || idx >= arr.size()) throw std::runtime_error("array index out of range");
    return arr[idx];
}

namespace {

struct Parser {
    const char* p;
    const char* end;
    const char* begin;
    std::string err;

    void fail(const char* why) {
        if (!err.empty()) return;
        int line = 1, col = 1;
        for (const char* q = begin; q < p && q < end; ++q) {
            if (*q == '\n') { ++line; col = 1; } else { ++col; }
        }
        char buf[160];
        std::snprintf(buf, sizeof buf, "line %d col %d: %s", line, col, why);
        err = buf;
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if ((size_t)(end - p) >= n && std::memcmp(p, w, n) == 0) { p += n; return true; }
        return false;
    }
    static void put_utf8(std::string& s, uint32_t c) {
        if (c < 0x80) s.push_back((char)c);
        else if (c < 0x800) { s.push_back((char)(0xC0 | (c >> 6))); s.push_back((char)(0x80 | (c & 0x3F))); }
        else if (c < 0x10000) {
            s.push_back((char)(0xE0 | (c >> 12))); s.push_back((char)(0x80 | ((c >> 6) & 0x3F)));
            s.push_back((char)(0x80 | (c & 0x3F)));
        } else {
            s.push_back((char)(0xF0 | (c >> 18))); s.push_back((char)(0x80 | ((c >> 12) & 0x3F)));
            s.push_back((char)(0x80 | ((c >> 6) & 0x3F))); s.push_back((char)(0x80 | (c & 0x3F)));
        }
    }
    bool hex4(uint32_t& v) {
        if (end - p < 4) { fail("truncated \\u escape"); return false; }
        v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else { fail("bad \\u escape"); return false; }
        }
        return true;
    }
    // A raw multi-byte UTF-8 sequence whose lead byte `c` was just read: checked
    // against the well-formed byte ranges of RFC 3629 (no overlong forms, no
    // encoded surrogates, nothing above U+10FFFF), as nlohmann-json's lexer
    // checks them, and appended to s.
    bool utf8_tail(unsigned char c, std::string& s) {
        int n = 0;
        unsigned char lo = 0x80, hi = 0xBF;  // range of the first continuation byte
        if (c >= 0xC2 && c <= 0xDF) n = 1;
        else if (c == 0xE0) { n = 2; lo = 0xA0; }
        else if ((c >= 0xE1 && c <= 0xEC) || c == 0xEE || c == 0xEF) n = 2;
        else if (c == 0xED) { n = 2; hi = 0x9F; }
        else if (c == 0xF0) { n = 3; lo = 0x90; }
        else if (c >= 0xF1 && c <= 0xF3) n = 3;
        else if (c == 0xF4) { n = 3; hi = 0x8F; }
        else { fail("ill-formed UTF-8 byte"); return false; }
        if (end - p < n) { fail("ill-formed UTF-8 byte"); return false; }
        for (int k = 0; k < n; ++k) {
            const unsigned char t = (unsigned char)p[k];
            if (t < (k == 0 ? lo : 0x80) || t > (k == 0 ? hi : 0xBF)) { fail("ill-formed UTF-8 byte"); return false; }
        }
        s.push_back((char)c);
        s.append(p, (size_t)n);
        p += n;
        return true;
    }
    bool string(std::string& s) {
        ++p;  // opening quote
        while (true) {
            if (p >= end) { fail("unterminated string"); return false; }
            char c = *p++;
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) { fail("control character in string"); return false; }
            if ((unsigned char)c >= 0x80) {
                if (!utf8_tail((unsigned char)c, s)) return false;
                continue;
            }
            if (c != '\\') { s.push_back(c); continue; }
            if (p >= end) { fail("unterminated escape"); return false; }
            char e = *p++;
            switch (e) {
                case '"': s.push_back('"'); break;
                case '\\': s.push_back('\\'); break;
                case '/': s.push_back('/'); break;
                case 'b': s.push_back('\b'); break;
                case 'f': s.push_back('\f'); break;
                case 'n': s.push_back('\n'); break;
                case 'r': s.push_back('\r'); break;
                case 't': s.push_back('\t'); break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xDC00 && cp <= 0xDFFF) { fail("lone low surrogate"); return false; }
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        uint32_t lo;
                        if (!(lit("\\u") && hex4(lo)) || lo < 0xDC00 || lo > 0xDFFF) {
                            fail("bad surrogate pair");
                            return false;
                        }
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(s, cp);
                    break;
                }
                default: fail("bad escape"); return false;
            }
        }
    }
    bool number(Value& v) {
        const char* s = p;
        bool real = false;
        if (p < end && *p == '-') ++p;
        if (p >= end || !(*p >= '0' && *p <= '9')) { fail("bad number"); return false; }
        if (*p == '0') ++p;
        else while (p < end && *p >= '0' && *p <= '9') ++p;
        if (p < end && *p == '.') {
            real = true; ++p;
            if (p >= end || !(*p >= '0' && *p <= '9')) { fail("bad fraction"); return false; }
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            real = true; ++p;
            if (p < end && (*p == '+' || *p == '-')) ++p;
            if (p >= end || !(*p >= '0' && *p <= '9')) { fail("bad exponent"); return false; }
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        std::string tok(s, p);
        if (!real) {
            // nlohmann: a negative literal as int64, a non-negative one as uint64,
            // and a double when it fits neither
            errno = 0;
            if (tok[0] == '-') {
                long long ll = std::strtoll(tok.c_str(), nullptr, 10);
                if (errno != ERANGE) { v.kind = Kind::Int; v.i = ll; v.d = (double)ll; return true; }
            } else {
                unsigned long long ull = std::strtoull(tok.c_str(), nullptr, 10);
                if (errno != ERANGE) {
                    if (ull <= (unsigned long long)INT64_MAX) { v.kind = Kind::Int; v.i = (int64_t)ull; }
                    else { v.kind = Kind::Uint; v.u = ull; }
                    v.d = (double)ull;
                    return true;
                }
            }
        }
        v.kind = Kind::Real;
        v.d = std::strtod(tok.c_str(), nullptr);  // correctly rounded in glibc
        // beyond binary64's range: nlohmann rejects it (out_of_range 406); an
        // underflow to zero or a denormal is kept
        if (!std::isfinite(v.d)) { p = s; fail("number overflow"); return false; }
        return true;
    }
    bool value(Value& v, int depth) {
        if (depth > 256) { fail("nesting too deep"); return false; }
        ws();
        if (p >= end) { fail("unexpected end of input"); return false; }
        char c = *p;
        if (c == '{') {
            ++p;
            v.kind = Kind::Object;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            while (true) {
                ws();
                if (p >= end || *p != '"') { fail("expected object key"); return false; }
                std::string key;
                if (!string(key)) return false;
                ws();
                if (p >= end || *p != ':') { fail("expected ':'"); return false; }
                ++p;
                Value child;
                if (!value(child, depth + 1)) return false;
                // a repeated key keeps its first position and takes the last
                // value, as nlohmann::ordered_json's parser does
                auto same = std::find_if(v.obj.begin(), v.obj.end(), [&](const auto& kv) { return kv.first == key; });
                if (same != v.obj.end()) same->second = std::move(child);
                else v.obj.emplace_back(std::move(key), std::move(child));
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                fail("expected ',' or '}'");
                return false;
            }
        }
        if (c == '[') {
            ++p;
            v.kind = Kind::Array;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            while (true) {
                Value child;
                if (!value(child, depth + 1)) return false;
                v.arr.push_back(std::move(child));
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                fail("expected ',' or ']'");
                return false;
            }
        }
        if (c == '"') { v.kind = Kind::String; return string(v.s); }
        if (lit("true")) { v.kind = Kind::Bool; v.b = true; return true; }
        if (lit("false")) { v.kind = Kind::Bool; v.b = false; return true; }
        if (lit("null")) { v.kind = Kind::Null; return true; }
        if (c == '-' || (c >= '0' && c <= '9')) return number(v);
        fail("unexpected character");
        return false;
    }
};

}  // namespace

bool parse(const char* text, size_t len, Value& out, std::string& err) {
    Parser ps{text, text + len, text, {}};
    // a leading UTF-8 byte order mark is skipped, as nlohmann-json does
    if (len >= 3 && std::memcmp(text, "\xEF\xBB\xBF", 3) == 0) ps.p += 3;
    out = Value{};
    if (!ps.value(out, 0)) { err = ps.err; return false; }
    ps.ws();
    if (ps.p != ps.end) { ps.fail("trailing characters after document"); err = ps.err; return false; }
    return true;
}

static std::runtime_error type_error(const char* what, const char* want) {
    return std::runtime_error(std::string("scene key '") + what + "' must be " + want);
}

double as_double(const Value& v, const char* what) {
    if (v.kind == Kind::Int) return (double)v.i;
    if (v.kind == Kind::Uint) return (double)v.u;
    if (v.kind == Kind::Real) return v.d;
    if (v.kind == Kind::Bool) return v.b ? 1.0 : 0.0;
    throw type_error(what, "a number");
}

float as_float(const Value& v, const char* what) {
    float f;
    if (v.kind == Kind::Int) f = static_cast<float>(v.i);
    else if (v.kind == Kind::Uint) f = static_cast<float>(v.u);  // directly, not through double
    else if (v.kind == Kind::Real) f = static_cast<float>(v.d);
    else if (v.kind == Kind::Bool) return v.b ? 1.0f : 0.0f;
    else throw type_error(what, "a number");
    // beyond binary32's range the reference's field would hold an infinity that
    // no later stage can use or save; such a value is an error here
    if (!std::isfinite(f)) throw type_error(what, "a number in float's range");
    return f;
}

int as_int(const Value& v, const char* what) {
    if (v.kind == Kind::Int) return static_cast<int>(v.i);
    if (v.kind == Kind::Uint) return static_cast<int>(v.u);
    if (v.kind == Kind::Real) {
        // nlohmann's get<int>() casts without a check (undefined beyond int's
        // range); such a value is an error here
        if (!(v.d > -2147483649.0 && v.d < 2147483648.0)) throw type_error(what, "a number in int's range");
        return static_cast<int>(v.d);
    }
    if (v.kind == Kind::Bool) return v.b ? 1 : 0;
    throw type_error(what, "a number");
}

bool as_bool(const Value& v, const char* what) {
    if (v.kind == Kind::Bool) return v.b;
    if (v.kind == Kind::Int) return v.i != 0;
    if (v.kind == Kind::Uint) return true;
    if (v.kind == Kind::Real) return v.d != 0.0;
    throw type_error(what, "a boolean");
}

const std::string& as_string(const Value& v, const char* what) {
    if (v.kind != Kind::String) throw type_error(what, "a string");
    return v.s;
}

// Strings as the reference saves them, nlohmann-json's dump(4, ' ', true)
// (src/pathtracer.cpp:3469): the two-character escapes for " \\ \b \f \n \r \t,
// and with ensure_ascii every other code point below 0x20 or from 0x7F up as
// \uXXXX in lower-case hex (a UTF-16 surrogate pair above U+FFFF). A byte that
// does not start a valid UTF-8 sequence is written as it is.
static void dump_string(std::string& o, const std::string& s) {
    auto u16 = [&](unsigned cp) {
        char b[16];
        std::snprintf(b, sizeof b, "\\u%04x", cp);
        o += b;
    };
    o.push_back('"');
    const size_t n = s.size();
    for (size_t i = 0; i < n; ++i) {
        const unsigned char c = (unsigned char)s[i];
        switch (c) {
            case '"': o += "\\\""; continue;
            case '\\': o += "\\\\"; continue;
            case '\n': o += "\\n"; continue;
            case '\r': o += "\\r"; continue;
            case '\t': o += "\\t"; continue;
            case '\b': o += "\\b"; continue;
            case '\f': o += "\\f"; continue;
            default: break;
        }
        if (c < 0x20 || c == 0x7F) { u16(c); continue; }
        if (c < 0x80) { o.push_back((char)c); continue; }
        // a multi-byte UTF-8 sequence: its length, lead bits and minimum value
        const int len = c >= 0xF0 && c < 0xF5 ? 4 : c >= 0xE0 ? (c < 0xF0 ? 3 : 0) : c >= 0xC2 ? 2 : 0;
        unsigned cp = len == 4 ? (c & 0x07u) : len == 3 ? (c & 0x0Fu) : (c & 0x1Fu);
        bool ok = len > 0 && i + (size_t)len <= n;
        for (int k = 1; ok && k < len; ++k) {
            const unsigned char t = (unsigned char)s[i + (size_t)k];
            ok = (t & 0xC0u) == 0x80u;
            cp = (cp << 6) | (t & 0x3Fu);
        }
        ok = ok && !(len == 3 && (cp < 0x800 || (cp >= 0xD800 && cp <= 0xDFFF))) && !(len == 4 && (cp < 0x10000 || cp > 0x10FFFF));
        if (!ok) { o.push_back((char)c); continue; }
        if (cp <= 0xFFFF) u16(cp);
        else { u16(0xD800 + ((cp - 0x10000) >> 10)); u16(0xDC00 + ((cp - 0x10000) & 0x3FF)); }
        i += (size_t)len - 1;
    }
    o.push_back('"');
}

// A double as nlohmann-json's serializer writes it (the reference saves scenes
// through it): the shortest digit string that reads back as the same value (as
// the same binary32 for `f32`), laid out as plain digits with ".0" when the
// decimal point falls after the last digit and within 15 places, as d.ddd when
// it falls inside, as 0.000ddd down to 10^-4, and otherwise as d.ddde+XX with a
// signed exponent of at least two digits; zero is 0.0 or -0.0.
static void dump_real(std::string& o, double x, bool f32) {
    if (std::signbit(x)) { o.push_back('-'); x = -x; }
    if (x == 0.0) { o += "0.0"; return; }
    char b[40];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(b, sizeof b, "%.*e", prec - 1, x);
        const double r = std::strtod(b, nullptr);
        if (f32 ? (float)r == (float)x : r == x) break;
    }
    // b is "d.ddde[+-]XX": value = 0.DIGITS x 10^n, n = XX + 1
    std::string dig;
    const char* q = b;
    for (; *q && *q != 'e'; ++q)
        if (*q >= '0' && *q <= '9') dig.push_back(*q);
    while (dig.size() > 1 && dig.back() == '0') dig.pop_back();
    const int k = (int)dig.size();
    const int n = std::atoi(q + 1) + 1;
    if (k <= n && n <= 15) {
        o += dig;
        o.append((size_t)(n - k), '0');
        o += ".0";
    } else if (0 < n && n <= 15) {
        o.append(dig, 0, (size_t)n);
        o.push_back('.');
        o.append(dig, (size_t)n, std::string::npos);
    } else if (-4 < n && n <= 0) {
        o += "0.";
        o.append((size_t)(-n), '0');
        o += dig;
    } else {
        o.push_back(dig[0]);
        if (k > 1) {
            o.push_back('.');
            o.append(dig, 1, std::string::npos);
        }
        std::snprintf(b, sizeof b, "e%c%02d", n - 1 < 0 ? '-' : '+', std::abs(n - 1));
        o += b;
    }
}

static void dump_rec(std::string& o, const Value& v, int indent, int level) {
    auto nl = [&](int lv) {
        o.push_back('\n');
        o.append((size_t)(indent * lv), ' ');
    };
    switch (v.kind) {
        case Kind::Null: o += "null"; break;
        case Kind::Bool: o += v.b ? "true" : "false"; break;
        case Kind::Int: o += std::to_string(v.i); break;
        case Kind::Uint: o += std::to_string(v.u); break;
        case Kind::Real:
            if (!std::isfinite(v.d)) o += "null";  // as nlohmann::json does
            else dump_real(o, v.d, v.f32);
            break;
        case Kind::String: dump_string(o, v.s); break;
        case Kind::Array:
            if (v.arr.empty()) { o += "[]"; break; }
            o.push_back('[');
            for (size_t k = 0; k < v.arr.size(); ++k) {
                nl(level + 1);
                dump_rec(o, v.arr[k], indent, level + 1);
                if (k + 1 < v.arr.size()) o.push_back(',');
            }
            nl(level);
            o.push_back(']');
            break;
        case Kind::Object:
            if (v.obj.empty()) { o += "{}"; break; }
            o.push_back('{');
            for (size_t k = 0; k < v.obj.size(); ++k) {
                nl(level + 1);
                dump_string(o, v.obj[k].first);
                o += ": ";
                dump_rec(o, v.obj[k].second, indent, level + 1);
                if (k + 1 < v.obj.size()) o.push_back(',');
            }
            nl(level);
            o.push_back('}');
            break;
    }
}

std::string dump(const Value& v, int indent) {
    std::string o;
    dump_rec(o, v, indent, 0);
    return o;
}

}  // namespace ptjson
