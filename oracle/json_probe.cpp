// Test infrastructure only. Parses a scene file with the JSON library the
// reference vendors (nlohmann-json 3.11.3, includes/json/json.hpp, used as
// nlohmann::ordered_json by src/pathtracer.cpp:1177) and prints what the
// reference's UpdateFromJSON (src/pathtracer.cpp:2576-2722) would read from each
// number and bool leaf. Those fields are assigned through the library's implicit
// conversions: float fields through get<float>() (binary64 parse, then a cast),
// int fields through get<int>().
//
// Output, one line per leaf: <JSON pointer> <f32 bits of get<float>(), hex>
// <get<int>(), or x when the value is outside int's range>.
//
// With --dump, prints instead the document as the reference's SaveScene writes
// it: scene.dump(4, ' ', true) (src/pathtracer.cpp:3465-3470).
//
// The default parse is ordered_json::parse of the whole text, as the startup
// scene is read (src/pathtracer.cpp:2573). With --stream the file is read as
// LoadScene reads one, `file >> json` (ReadJSON, src/pathtracer.cpp:893-908),
// which parses the first value and leaves what follows it unread.
//
// Built by oracle/Makefile into oracle/_ref/ from the header where it lies in
// the reference tree; never part of the product.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include <json/json.hpp>

using Json = nlohmann::ordered_json;

static void leaf(const std::string& ptr, const Json& j) {
    const float f = j.get<float>();
    uint32_t bits;
    std::memcpy(&bits, &f, 4);
    std::printf("%s %08x ", ptr.c_str(), bits);
    bool in_range = false;
    if (j.is_number_unsigned()) in_range = j.get<uint64_t>() <= (uint64_t)INT32_MAX;
    else if (j.is_number_integer()) in_range = j.get<int64_t>() >= INT32_MIN && j.get<int64_t>() <= INT32_MAX;
    else if (j.is_number_float()) in_range = std::fabs(j.get<double>()) < 2147483647.0;
    else in_range = j.is_boolean();
    if (in_range) std::printf("%d\n", j.get<int>());
    else std::printf("x\n");
}

static void walk(const std::string& ptr, const Json& j) {
    if (j.is_object()) {
        for (auto it = j.begin(); it != j.end(); ++it) walk(ptr + "/" + it.key(), it.value());
    } else if (j.is_array()) {
        for (size_t i = 0; i < j.size(); ++i) walk(ptr + "/" + std::to_string(i), j[i]);
    } else if (j.is_number() || j.is_boolean()) {
        leaf(ptr, j);
    }
}

int main(int argc, char** argv) {
    const bool dump = argc == 3 && std::strcmp(argv[1], "--dump") == 0;
    const bool stream = argc == 3 && std::strcmp(argv[1], "--stream") == 0;
    if (argc != 2 && !dump && !stream) {
        std::fprintf(stderr, "usage: %s [--dump | --stream] scene.json\n", argv[0]);
        return 2;
    }
    const char* path = argv[argc - 1];
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", path);
        return 2;
    }
    try {
        if (stream) {
            Json j;
            in >> j;
            walk("", j);
            return 0;
        }
        std::stringstream ss;
        ss << in.rdbuf();
        const Json j = Json::parse(ss.str());
        if (dump) std::fputs(j.dump(4, ' ', true).c_str(), stdout);
        else walk("", j);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "parse error: %s\n", e.what());
        return 1;
    }
    return 0;
}
