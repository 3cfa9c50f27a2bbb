/* Checks the division of the fused FP64 step (div_rn: x / b from the
 * correctly rounded reciprocal 1/b and two FMA residual corrections) against
 * IEEE division, bit for bit, for every finite quotient: random significands
 * and exponents, random and fixed divisors of both signs, the neighbourhood of
 * each fixed divisor, and signed zeros and extreme values.  Exit status 0 when
 * every quotient matches.  Usage: div_rn_check [random cases] */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double div_rn(double x, double b, double rb) {
    const double q0 = x * rb;
    const double q1 = fma(fma(-q0, b, x), rb, q0);
    return copysign(fma(fma(-q1, b, x), rb, q1), q0);
}
static uint64_t bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static double from_bits(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint64_t next(uint64_t* s) { *s ^= *s << 13; *s ^= *s >> 7; *s ^= *s << 17; return *s; }

#define NFIX 24
static double fixed[NFIX];

static int differs(double x, double b) {
    const double want = x / b, got = div_rn(x, b, 1.0 / b);
    return isfinite(want) && bits(want) != bits(got);
}

struct job { int id; long n; long bad; };
static void* run(void* p) {
    struct job* j = (struct job*)p;
    uint64_t s = 0x9e3779b97f4a7c15ull * (uint64_t)(j->id + 1);
    for (long k = 0; k < j->n; ++k) {
        double b;
        if ((k & 3) == 0) b = from_bits(((uint64_t)(1023 + (int)(next(&s) % 121) - 60) << 52) | (next(&s) >> 12));
        else b = fixed[next(&s) % NFIX];
        if (next(&s) & 1) b = -b;
        double x = from_bits(((uint64_t)(1023 + (int)(next(&s) % 401) - 200) << 52) | (next(&s) >> 12));
        if (next(&s) & 1) x = -x;
        if (differs(x, b)) {
            if (j->bad < 3) printf("mismatch: x=%a b=%a\n", x, b);
            j->bad++;
        }
    }
    return NULL;
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 100000000L;
    const double pi = 3.14159265358979323846;
    const double f[NFIX] = {2 * pi / 256, 2 * pi / 64, 2 * pi / 16, 2 * pi / 512, 2 * pi / 10, 0.1, 20.0 / 200,
                            6.0 / 60, 1.0, 100.0, 10.0, 0.3, 3.0, 7.0, 0.5, 1.0 / 3,
                            (2 * pi / 256) * (2 * pi / 256) * (2 * pi / 256), 0.1 * 0.1 * 0.1, 6.693874020129442e-4,
                            1.447867e-2, 0.999999999999, 1.0000000000001, 1.9999999999999998, 1e-3};
    memcpy(fixed, f, sizeof f);
    long bad = 0;
    const double special[] = {0.0, 1.0, 2.0, 0.5, 3.0, 1e-300, 1e300, 2.2250738585072014e-308,
                              1.7976931348623157e308, 1.0000000000000002, 0.9999999999999999};
    for (int i = 0; i < NFIX; ++i)
        for (unsigned k = 0; k < sizeof special / sizeof special[0]; ++k)
            for (int sb = 0; sb < 2; ++sb)
                for (int sx = 0; sx < 2; ++sx) {
                    const double b = sb ? -fixed[i] : fixed[i], x = sx ? -special[k] : special[k];
                    if (differs(x, b)) { printf("special mismatch: x=%a b=%a\n", x, b); bad++; }
                }
    for (int i = 0; i < NFIX; ++i)
        for (long d = -200000; d <= 200000; ++d) {
            const double x = from_bits(bits(fixed[i]) + (uint64_t)d);
            if (differs(x, fixed[i])) { if (bad < 3) printf("mismatch near b: x=%a b=%a\n", x, fixed[i]); bad++; }
        }
    enum { T = 8 };
    pthread_t th[T];
    struct job jobs[T];
    for (int t = 0; t < T; ++t) {
        jobs[t].id = t; jobs[t].n = n / T; jobs[t].bad = 0;
        pthread_create(&th[t], NULL, run, &jobs[t]);
    }
    for (int t = 0; t < T; ++t) { pthread_join(th[t], NULL); bad += jobs[t].bad; }
    printf("div_rn: %ld mismatches against IEEE division\n", bad);
    return bad != 0;
}
