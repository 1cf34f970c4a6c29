// Mutation fuzzing of the scene JSON reader and writer (pathtracer_amd/csrc/json.cpp)
// and of the scene loader, packer, writer and SDF recogniser built on them
// (pathtracer_amd/csrc/scene.cpp), on the CPU, meant to be built with
// AddressSanitizer and UndefinedBehaviorSanitizer. Each seed file is mutated
// (byte flips, inserted special bytes and tokens, truncation, duplicated or
// deleted spans) and parsed. For every document that parses, the writer's text
// must parse again and write the same text (a fixed point), and the numeric
// conversions run on every number. Every document that loads as a scene is
// packed, its camera taken, and written; the written scene must load again and
// write the same text. The seeds' SDF GLSL texts are mutated and recognised.
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
#include "scene.hpp"

static void touch_numbers(const ptjson::Value& v, double& sink) {
    using ptjson::Kind;
    if (v.is_number() || v.kind == Kind::Bool) {
        // values outside float's or int's range are errors (exceptions), never undefined casts
        try { sink += (double)ptjson::as_float(v, "x"); } catch (const std::exception&) {}
        try { sink += ptjson::as_int(v, "x"); } catch (const std::exception&) {}
        sink += ptjson::as_double(v, "x") * 0.0;
        sink += ptjson::as_bool(v, "x") ? 1.0 : 0.0;
    }
    for (const auto& c : v.arr) touch_numbers(c, sink);
    for (const auto& kv : v.obj) touch_numbers(kv.second, sink);
}

static long loaded = 0;

// A document that loads as a scene: packed, its first camera shot taken, written;
// the written text loads again and writes the same text.
static bool scene_round_trip(const std::string& text, int it) {
    pt::SceneModel m;
    std::string err;
    try {
        if (pt::load_scene_json(text.data(), text.size(), m, err) != PT_OK) return true;
    } catch (const std::exception&) {
        return true;  // the C ABI turns these into PT_ERR_JSON
    }
    ++loaded;
    pt_packed_scene ubo;
    pt_push_constants pc{};
    (void)pt::pack_reference_ubo(m, ubo, err);
    (void)pt::camera_to_push(m, 1, pc, err);
    const std::string once = pt::write_scene_json(m);
    pt::SceneModel m2;
    if (pt::load_scene_json(once.data(), once.size(), m2, err) != PT_OK) {
        std::printf("FAIL written scene does not load (iteration %d): %s\n", it, err.c_str());
        return false;
    }
    if (pt::write_scene_json(m2) != once) {
        std::printf("FAIL scene writing is not a fixed point (iteration %d)\n", it);
        return false;
    }
    return true;
}

// The seeds' SDF texts, mutated, through the recogniser.
static bool fuzz_glsl(const std::vector<std::string>& seeds, int iters, std::mt19937& rng, long& recognised) {
    std::vector<std::string> texts;
    for (const std::string& s : seeds) {
        ptjson::Value v;
        std::string err;
        if (!ptjson::parse(s.data(), s.size(), v, err)) continue;
        if (const ptjson::Value* arr = v.get("sdf"))
            for (const auto& e : arr->arr)
                if (const ptjson::Value* g = e.get("glsl")) texts.push_back(g->s);
    }
    if (texts.empty()) return true;
    static const char* toks[] = {"(", ")", "{", "}", ";", "for", "return", "1e39", "-", "0.0", "i < 100000",
                                 "p.x", "/*", "//", "\n", "1.", ".5"};
    for (int it = 0; it < iters; ++it) {
        std::string g = texts[(size_t)it % texts.size()];
        for (int e = 0, n = 1 + (int)(rng() % 3); e < n && !g.empty(); ++e) {
            const size_t pos = rng() % (g.size() + 1);
            switch (rng() % 4) {
                case 0: if (pos < g.size()) g[pos] = (char)(rng() & 0xFF); break;
                case 1: g.insert(pos, toks[rng() % (sizeof toks / sizeof toks[0])]); break;
                case 2: g.erase(pos, rng() % 16); break;
                default: if (pos < g.size()) g[pos] = "0123456789.-+*/()<>;{} \n"[rng() % 24]; break;
            }
        }
        pt_sdf_desc d;
        std::string err;
        try {
            if (pt::recognize_sdf(g, d, err) == PT_OK) ++recognised;
            else if (err.empty()) {
                std::printf("FAIL recogniser rejected without a message (iteration %d)\n", it);
                return false;
            }
        } catch (const std::exception&) {
        }
    }
    return true;
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
        if (!scene_round_trip(s, it)) return 1;
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
    long recognised = 0;
    if (!fuzz_glsl(seeds, iters, rng, recognised)) return 1;
    std::printf("ok parsed=%ld rejected=%ld loaded=%ld recognised=%ld sink=%d\n", parsed, rejected, loaded, recognised,
                sink != sink ? 1 : 0);
    return 0;
}
