// Accuracy of emx_exp / emx_log (ember_amd/csrc/emx_math.h) against long double expl/logl
// on the host: prints the largest error in units of the last place over seeded samples
// of the argument ranges the kinetics use, and the special cases.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "emx_math.h"

static double ulp_err(double got, long double ref) {
    if (std::isinf((double)ref) && got == (double)ref) return 0.0;
    const double r = (double)ref;
    if (r == 0.0) return got == 0.0 ? 0.0 : 1e300;
    int e;
    std::frexp(r, &e);
    const double ulp = std::ldexp(1.0, std::max(e - 53, -1074));
    return (double)(std::fabs((long double)got - ref) / ulp);
}

int main() {
    std::mt19937_64 rng(7);
    double max_exp = 0.0, max_log = 0.0, worst_exp = 0.0, worst_log = 0.0;
    // exp: Arrhenius exponents, Gibbs factors and Troe terms span about [-745, 709]
    std::uniform_real_distribution<double> ux(-745.0, 709.0), us(-2.0, 2.0);
    for (int i = 0; i < 2000000; i++) {
        const double x = (i & 1) ? ux(rng) : us(rng);
        const double e = ulp_err(emx::emx_exp(x), expl((long double)x));
        if (e > max_exp) { max_exp = e; worst_exp = x; }
    }
    // log: temperatures, pressure ratios, falloff arguments down to 1e-300
    std::uniform_real_distribution<double> ul(-300.0, 300.0), un(0.5, 2.0);
    for (int i = 0; i < 2000000; i++) {
        const double x = (i & 1) ? std::pow(10.0, ul(rng)) : un(rng);
        const double e = ulp_err(emx::emx_log(x), logl((long double)x));
        if (e > max_log) { max_log = e; worst_log = x; }
    }
    bool special = emx::emx_exp(0.0) == 1.0 && emx::emx_log(1.0) == 0.0 &&
                   emx::emx_exp(-800.0) == 0.0 && std::isinf(emx::emx_exp(800.0)) &&
                   emx::emx_exp(NAN) == 0.0 &&  // documented: the clamps map NaN to the lower limit
                   std::isinf(emx::emx_exp(INFINITY)) && emx::emx_exp(-INFINITY) == 0.0 && emx::emx_exp(-1e30) == 0.0 &&
                   std::isinf(emx::emx_exp(1e30)) && emx::emx_log(2.0) == (double)logl(2.0L);
    // denormal results and arguments
    const double ed = ulp_err(emx::emx_exp(-720.0), expl(-720.0L));
    const double ld = ulp_err(emx::emx_log(1e-310), logl(1e-310L));
    std::printf("exp_max_ulp %.3f at %.17g\nlog_max_ulp %.3f at %.17g\nexp_denormal_ulp %.3f\n"
                "log_denormal_ulp %.3f\nspecial %d\n",
                max_exp, worst_exp, max_log, worst_log, ed, ld, special ? 1 : 0);
    return 0;
}
