// Host code of libhipdvr.so under AddressSanitizer and UndefinedBehavior-
// Sanitizer: the .vvd parser (io.cpp) on the fixture files and on damaged
// copies of them, the TF table and pre-integration builders (tf.cpp) on
// random keys and edge cases.  Built by
// tests/test_sanitizers.py with -fsanitize=address,undefined
// -fno-sanitize-recover=all; any finding aborts with a report.  Argument: the
// directory holding the fixture .vvd/.raw files and a scratch directory.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hipdvr.h"
#include "hipdvr_io.h"

static uint64_t rs = 0x1234567ull;
static uint32_t rnd(uint32_t n) {
    rs = rs * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(rs >> 33) % n;
}

static std::string slurp(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    std::stringstream s;
    s << f.rdbuf();
    return s.str();
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string fix = argv[1], tmp = argv[2];
    long checks = 0;
    // the parser on the fixtures and on damaged copies: truncated at every
    // length step, bytes overwritten, attributes emptied
    for (const char* name : {"nucleon.vvd", "walnut.vvd"}) {
        const std::string text = slurp(fix + "/" + name);
        hipdvr_volume_desc d;
        char raw[512];
        if (hipdvr_vvd_parse((fix + "/" + name).c_str(), &d, raw, sizeof raw) != HIPDVR_OK) {
            std::printf("fixture %s did not parse\n", name);
            return 1;
        }
        for (size_t cut = 0; cut < text.size(); cut += 7) {
            const std::string p = tmp + "/cut.vvd";
            std::ofstream(p, std::ios::binary) << text.substr(0, cut);
            (void)hipdvr_vvd_parse(p.c_str(), &d, raw, sizeof raw);
            (void)hipdvr_vvd_parse(p.c_str(), &d, raw, 8);     // short path buffer
            ++checks;
        }
        for (int k = 0; k < 400; ++k) {
            std::string t = text;
            const int n = 1 + (int)rnd(4);
            for (int j = 0; j < n; ++j) t[rnd((uint32_t)t.size())] = "<>\"=/ x0-9\n\0"[rnd(12)];
            const std::string p = tmp + "/bad.vvd";
            std::ofstream(p, std::ios::binary) << t;
            (void)hipdvr_vvd_parse(p.c_str(), &d, raw, sizeof raw);
            ++checks;
        }
    }
    (void)hipdvr_vvd_parse((tmp + "/missing.vvd").c_str(), nullptr, nullptr, 0);
    // TF tables and pre-integration tables
    std::vector<uint8_t> tf(4 * 4096);
    std::vector<float> pre;
    for (int k = 0; k < 300; ++k) {
        const int nk = (int)rnd(9);
        std::vector<hipdvr_tf_key> keys((size_t)(nk ? nk : 1));
        float x = 0.f;
        for (int i = 0; i < nk; ++i) {
            x += (float)rnd(1000) / 4000.f;
            keys[(size_t)i].intensity = x;
            keys[(size_t)i].split = (int32_t)rnd(2);
            for (int c = 0; c < 4; ++c) {
                keys[(size_t)i].color_l[c] = (uint8_t)rnd(256);
                keys[(size_t)i].color_r[c] = (uint8_t)rnd(256);
            }
        }
        const int width = 1 + (int)rnd(4096);
        const float gamma = rnd(3) ? 1.f : (float)rnd(400) / 100.f;
        const float lo = (float)rnd(100) / 200.f, hi = lo + (float)rnd(100) / 100.f;
        (void)hipdvr_tf_from_keys(keys.data(), nk, width, gamma, lo, hi, (int32_t)rnd(4) - 1, tf.data());
        const uint32_t n = 1 + rnd(96);
        pre.assign(4 * (size_t)n * n, 0.f);
        (void)hipdvr_preint_build(keys.data(), nk, gamma, lo, hi, (int32_t)rnd(3), n,
                                  (float)(1 + rnd(100)) / 1000.f, (int32_t)rnd(2), pre.data());
        checks += 2;
    }
    (void)hipdvr_tf_from_keys(nullptr, 0, 256, 1.f, 0.f, 1.f, 1, tf.data());
    (void)hipdvr_preint_build(nullptr, 0, 1.f, 0.f, 1.f, 1, 64, 0.01f, 0, nullptr);
    std::printf("ok %ld\n", checks);
    return 0;
}
