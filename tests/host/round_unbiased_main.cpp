// Host-side check of the mixed linear step's unbiased FP64 -> FP32 rounding
// (round_unbiased / mix_bits in gs_dev.h, compiled here for the CPU):
// results are always one of the two FP32 neighbours of x, exact values and
// IEEE specials pass unchanged, and the mean rounding error over hashed draws
// is zero to sampling accuracy at several positions between neighbours and
// both signs (round to nearest would give errors of q and q - 1 ulp).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "gs_dev.h"

static int fails = 0;
static void check(bool ok, const char* what, double a = 0, double b = 0)
{
    if (!ok) {
        std::printf("FAIL %s (%.9g, %.9g)\n", what, a, b);
        ++fails;
    }
}

int main()
{
    const int n = 400000;
    const double qs[] = {0.05, 0.25, 0.5, 0.75, 0.95};
    const double bases[] = {1.0, 3.0e-7, 7.5e12};
    for (double base : bases)
        for (int sg = -1; sg <= 1; sg += 2)
            for (double q : qs) {
                const float lo = (float)base;                       // exact in FP32 for these bases
                const float hi = std::nextafterf(lo, INFINITY);
                const double ulp = (double)hi - (double)lo;
                const double x = sg * ((double)lo + q * ulp);
                double sum = 0;
                for (int i = 0; i < n; ++i) {
                    const float f = gs::round_unbiased(x, gs::mix_bits((unsigned)i * 0x9e3779b1u));
                    check(f == (float)(sg * lo) || f == (float)(sg * hi), "not a neighbour", x, f);
                    sum += (double)f;
                }
                const double bias = (sum / n - x) / ulp;
                std::printf("base %.3g sign %+d q %.2f: mean error %+.4f ulp\n", base, sg, q, bias);
                check(std::fabs(bias) < 0.004, "biased", x, bias);
            }
    for (unsigned u : {0u, 1u, 0x1fffffffu, 0xffffffffu}) {
        check(gs::round_unbiased(1.5, u) == 1.5f, "exact value moved", 1.5, u);
        check(gs::round_unbiased(-2.0, u) == -2.0f, "exact value moved", -2.0, u);
        check(gs::round_unbiased(0.0, u) == 0.0f && !std::signbit(gs::round_unbiased(0.0, u)), "+0", 0, u);
        check(gs::round_unbiased(-0.0, u) == 0.0f && std::signbit(gs::round_unbiased(-0.0, u)), "-0", 0, u);
        check(std::isinf(gs::round_unbiased(INFINITY, u)) && gs::round_unbiased(INFINITY, u) > 0, "+inf", 0, u);
        check(std::isinf(gs::round_unbiased(-INFINITY, u)) && gs::round_unbiased(-INFINITY, u) < 0, "-inf", 0, u);
        check(std::isnan(gs::round_unbiased(NAN, u)), "nan", 0, u);
        check(!std::isnan(gs::round_unbiased(3.4028234663852886e38, u)), "FLT_MAX", 0, u);
    }
    // the hash's low 29 bits, as used, are uniform: mean and the fraction below a quarter
    double m = 0, quarter = 0;
    for (int i = 0; i < n; ++i) {
        const unsigned v = gs::mix_bits((unsigned)i) & 0x1fffffffu;
        m += v / 536870912.0;
        quarter += v < 134217728u;
    }
    std::printf("hash: mean %.4f, below 1/4 %.4f\n", m / n, quarter / n);
    check(std::fabs(m / n - 0.5) < 0.003 && std::fabs(quarter / n - 0.25) < 0.003, "hash not uniform", m / n, quarter / n);
    std::printf(fails ? "FAILED %d\n" : "ok\n", fails);
    return fails ? 1 : 0;
}
