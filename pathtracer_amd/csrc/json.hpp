// json.hpp — minimal JSON document reader for scene files.
//
// The reference reads scenes with nlohmann::ordered_json 3.11.3
// (src/pathtracer.cpp:893-908) and converts numbers with implicit
// `float x = json` (src/pathtracer.cpp:2576-2722). nlohmann keeps a negative
// integer literal as int64, a non-negative one as uint64, one that fits neither
// and a literal with '.' or an exponent as double, and get<float>() is
// static_cast<float> of whichever it holds. This reader keeps the same
// distinction so the floats it yields are the same: an integer above 2^63
// converted through double can round differently from the direct conversion.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace ptjson {

// Int: an integer literal that fits int64; Uint: a non-negative one above INT64_MAX.
enum class Kind { Null, Bool, Int, Uint, Real, String, Array, Object };

struct Value {
    Kind kind = Kind::Null;
    bool b = false;
    int64_t i = 0;
    uint64_t u = 0;
    double d = 0.0;
    bool f32 = false;  // Real holding a binary32 value: written as the shortest text that re-reads as that float
    std::string s;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;  // insertion order, like ordered_json

    bool is_number() const { return kind == Kind::Int || kind == Kind::Uint || kind == Kind::Real; }
    // Member lookup; nullptr when absent or when this is not an object.
    const Value* get(const std::string& key) const;
    // Number of elements (array or object), 0 for null (missing keys read as null).
    size_t size() const;
    const Value& at(size_t idx) const;
};

// Parses `text`; on failure returns false and sets `err` to "line L col C: why".
bool parse(const char* text, size_t len, Value& out, std::string& err);

// Conversions with nlohmann's semantics; throw std::runtime_error on type mismatch.
float as_float(const Value& v, const char* what);
double as_double(const Value& v, const char* what);
int as_int(const Value& v, const char* what);
bool as_bool(const Value& v, const char* what);
const std::string& as_string(const Value& v, const char* what);

// Writes `v` as indented JSON (used for round-trip tests and scene export).
std::string dump(const Value& v, int indent = 4);

}  // namespace ptjson
