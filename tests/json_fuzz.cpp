// Mutation fuzzing of the scene JSON reader and writer (pathtracer_amd/csrc/json.cpp)
// on the CPU, meant to be built with AddressSanitizer and UndefinedBehaviorSanitizer.
// Each seed file is mutated (byte flips, inserted special bytes and tokens,
// truncation, duplicated or deleted spans) and parsed. For every document that
// parses, the writer's text must parse again and write the same text (a fixed
// point), and the numeric conversions run on every number.
//
// Usage: json_fuzz iterations seed file...
// Prints "ok parsed=P rejected=R" or the first failure, and exits 0 or 1.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "json.hpp"

static void touch_numbers(const ptjson::Value& v, double& sink) {
    using ptjson::Kind;
    if (v.is_number() || v.kind == Kind::Bool) {
        sink += (double)ptjson::as_float(v, "x");
        if (v.kind != Kind::Real || (v.d > -2e9 && v.d < 2e9)) sink += ptjson::as_int(v, "x");
        sink += ptjson::as_double(v, "x") * 0.0;
        sink += ptjson::as_bool(v, "x") ? 1.0 : 0.0;
    }
    for (const auto& c : v.arr) touch_numbers(c, sink);
    for (const auto& kv : v.obj) touch_numbers(kv.second, sink);
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s iterations seed file...\n", argv[0]);
        return 2;
    }
    const int iters = std::atoi(argv[1]);
    std::mt19937 rng((unsigned)std::atoi(argv[2]));
    std::vector<std::string> seeds;
    for (int i = 3; i < argc; ++i) {
        std::ifstream f(argv[i], std::ios::binary);
        std::stringstream ss;
        ss << f.rdbuf();
        seeds.push_back(ss.str());
    }
    static const char* tokens[] = {"\"", "\\", "\\u", "\\ud800", "\\udc00", "\\u00e9", "{", "}", "[", "]", ",", ":",
                                   "-", "1e400", "1e-400", "0.", "01", "18446744073709551616", "null", "true",
                                   "\xef\xbb\xbf", "\xc0\x80", "\xed\xa0\x80", "\xf4\x90\x80\x80", "\xe2\x82",
                                   "\x7f", "\x01", "[[[[[[[[", "]]]]]]]]", "\"a\":1,"};
    const int ntok = (int)(sizeof tokens / sizeof tokens[0]);
    long parsed = 0, rejected = 0;
    double sink = 0.0;
    for (int it = 0; it < iters; ++it) {
        std::string s = seeds[(size_t)(it % (int)seeds.size())];
        const int edits = 1 + (int)(rng() % 4);
        for (int e = 0; e < edits && !s.empty(); ++e) {
            const size_t pos = rng() % (s.size() + 1);
            switch (rng() % 6) {
                case 0: if (pos < s.size()) s[pos] = (char)(rng() & 0xFF); break;
                case 1: s.insert(pos, tokens[rng() % (unsigned)ntok]); break;
                case 2: s.resize(pos); break;
                case 3: {
                    const size_t len = rng() % 64;
                    if (pos < s.size()) s.insert(pos, s.substr(pos, len));
                    break;
                }
                case 4: {
                    const size_t len = rng() % 64;
                    if (pos < s.size()) s.erase(pos, len);
                    break;
                }
                default: if (pos < s.size()) s[pos] = "0123456789.eE-+\"{}[],: \n"[rng() % 25]; break;
            }
        }
        ptjson::Value v;
        std::string err;
        if (!ptjson::parse(s.data(), s.size(), v, err)) {
            if (err.empty()) {
                std::printf("FAIL rejected without a message (iteration %d)\n", it);
                return 1;
            }
            ++rejected;
            continue;
        }
        ++parsed;
        touch_numbers(v, sink);
        const std::string once = ptjson::dump(v);
        ptjson::Value v2;
        if (!ptjson::parse(once.data(), once.size(), v2, err)) {
            std::printf("FAIL written text does not parse (iteration %d): %s\n", it, err.c_str());
            return 1;
        }
        const std::string twice = ptjson::dump(v2);
        if (twice != once) {
            std::printf("FAIL writing is not a fixed point (iteration %d)\n", it);
            return 1;
        }
    }
    std::printf("ok parsed=%ld rejected=%ld sink=%d\n", parsed, rejected, sink != sink ? 1 : 0);
    return 0;
}
