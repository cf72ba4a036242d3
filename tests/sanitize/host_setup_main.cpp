// Sanitizer driver for the host set-up of gsolver_amd (config lookup, grids,
// dispersion, coefficient tables, Keldysh table, QDHT plan, dump parsing,
// initial field): tests/test_sanitizers.py builds it with gs_host.cpp and
// gs_special.cpp under -fsanitize=address,undefined and runs it on a config.
// Touches every table, then checks that a missing key raises gs::Error.
#include <cstdio>
#include <string>
#include <vector>

#include "gs_host.h"

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::vector<std::string> ov(argv + 2, argv + argc);
    const gs::Config cfg(argv[1], ov);
    const gs::Setup s = gs::build_setup(cfg, "");
    double acc = 0;
    for (double v : s.omega) acc += v;
    for (double v : s.ho) acc += v;
    for (double v : s.kerr) acc += v;
    for (double v : s.ionrate_ln) acc += v;
    for (double v : s.qx) acc += v;
    for (double v : s.qm1) acc += v;
    for (double v : s.qC) acc += v;
    for (double v : s.ZNET) acc += v;
    for (const auto& v : s.wavenum) acc += std::abs(v);
    for (const auto& v : s.plasma_func) acc += std::abs(v);
    for (const auto& v : s.raman_function) acc += std::abs(v);
    for (const auto& v : s.field_t) acc += std::abs(v);
    for (int r = 0; r < s.NR; ++r)
        for (int w = 0; w < s.NT; w += 7) acc += std::abs(gs::init_spectral_factor(s, r, w));
    std::printf("setup %d x %d, checksum %.17g\n", s.NR, s.NT, acc);
    try {
        (void)cfg.f("NO_SUCH_KEY_IN_ANY_CONFIG");
        std::printf("missing key did not raise\n");
        return 3;
    } catch (const gs::Error& e) {
        std::printf("missing key: %s\n", e.what());
    }
    return 0;
}
