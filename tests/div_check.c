/* div_check.c -- the FP64 kernels' division a / b as q = RN(a y), r = fma(-q, b, a),
 * RN(q + r y) with y = RN(1/b) (Markstein's correction, fused_common.hpp div_cr) against IEEE
 * division, on seeded random dividends of every sign and of exponents -60..4 and on divisors 3
 * and the relaxation times of the workloads.  Built and run by tests/test_div_cr.py.  Prints the
 * number of mismatches; exit status 0 when there are none. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t s = 88172645463325252ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    const double bs[] = {3.0, 0.5 + 3 * 0.05 * 20 / 100.0, 0.5 + 3 * 0.05 * 24 / 200.0, 0.5 + 3 * 0.05 * 40 / 300.0,
                         0.5 + 0.1 * 8.0 / 100.0 / (1.0 / 3.0), 0.5 + 0.1 * 12 / 100.0 / (1.0 / 3.0), 0.56, 0.5001, 0.8};
    long bad = 0;
    for (unsigned k = 0; k < sizeof(bs) / sizeof(bs[0]); k++) {
        const double b = bs[k], y = 1.0 / b;
        for (long t = 0; t < n; t++) {
            const uint64_t m = next();
            const int e = (int)(next() % 65) - 60;
            const double a = ldexp(1.0 + (double)(m >> 12) * 0x1p-52, e) * ((m & 1) ? -1.0 : 1.0);
            const double q = a * y, r = fma(-q, b, a), q2 = fma(r, y, q);
            if (q2 != a / b) bad++;
        }
    }
    printf("%ld\n", bad);
    return bad != 0;
}
