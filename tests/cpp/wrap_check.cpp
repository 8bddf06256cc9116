// Host-side check of the position wrap of the integrator epilogues
// (mdx::wrap1, cube::correct_position_and_momentum of geometry.h:403-445):
// fed NaN, +-inf and huge values it must return (no unbounded loop), report
// the non-finite inputs, and leave every finite coordinate inside [0, l];
// one box length out it must match the reference's single subtraction.
// Prints one line per case and "ok" at the end; exit status 1 on a failure.
#include "mdx_device.h"

#include <cmath>
#include <cstdio>
#include <limits>

using mdx::wrap1;

static int fails = 0;

static void expect(bool c, const char *what, double x)
{
    std::printf("%-48s x=%.17g %s\n", what, x, c ? "ok" : "FAIL");
    if (!c)
        fails++;
}

int main()
{
    const double l = 27.783;
    const double inf = std::numeric_limits<double>::infinity();
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int periodic = 0; periodic < 2; periodic++) {
        std::printf("periodic=%d\n", periodic);
        // non-finite coordinate or momentum: rejected, untouched
        const double bad[] = {nan, inf, -inf};
        for (double b : bad) {
            double x = b;
            double p = 1.0;
            const bool r = wrap1(x, p, l, periodic);
            expect(!r && !(x == x) == !(b == b) && p == 1.0, "non-finite x rejected", x);
        }
        {
            double x = 3.0;
            float p = std::numeric_limits<float>::infinity();
            expect(!wrap1(x, p, l, periodic) && x == 3.0, "non-finite p rejected", x);
        }
        // huge finite values: bounded, inside the box
        const double huge[] = {1e300, -1e300, 1e30, -1e30, 1e6 * l + 0.3, -1e6 * l - 0.3,
                               std::numeric_limits<double>::max(),
                               -std::numeric_limits<double>::max()};
        for (double h : huge) {
            double x = h;
            double p = 0.5;
            const bool r = wrap1(x, p, l, periodic);
            expect(r && std::isfinite(x) && x >= 0 && x <= l, "huge x brought inside", x);
        }
        // the reference's single shift / reflection one box length out
        const double near[] = {-1e-12, -0.3, l + 0.2, l + 1e-9, 0.0, l, 13.0};
        for (double v : near) {
            double x = v, p = 0.7;
            wrap1(x, p, l, periodic);
            double ref = v, pref = 0.7;
            if (periodic) {
                if (ref < 0)
                    while (ref < 0)
                        ref += l;
                else if (ref > l)
                    while (ref > l)
                        ref -= l;
            } else if (ref < 0) {
                ref = -ref;
                pref = -pref;
            } else if (ref > l) {
                ref = l - (ref - l);
                pref = -pref;
            }
            expect(x == ref && p == pref, "one box length out: reference rounding", x);
        }
        // several box lengths out (within the bound): the reference's loop
        {
            double x = -5.5 * l, p = 1.0, ref = -5.5 * l;
            wrap1(x, p, l, 1);
            while (ref < 0)
                ref += l;
            if (periodic)
                expect(x == ref, "five box lengths out: reference loop", x);
        }
    }
    std::printf(fails ? "FAILED %d\n" : "ok\n", fails);
    return fails ? 1 : 0;
}
