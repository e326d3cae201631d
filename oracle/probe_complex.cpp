// Pins how g++ evaluates std::complex<float> arithmetic, which the reference's
// FresnelComplex (include/bulbit/scattering.h:102-121) relies on.  For seeded
// random operands it counts how often each operation equals an explicit
// formula; the oracle restates exactly the formulas that match everywhere.
// Test infrastructure for the oracle.
#include <cmath>
#include <complex>
#include <cstdio>

static float rnd(unsigned* s)
{
    *s = *s * 1664525u + 1013904223u;
    return ((*s >> 8) * (1.0f / 16777216.0f)) * 4.0f - 2.0f;
}

int main()
{
    unsigned s = 12345u;
    const int n = 200000;
    int div_double = 0, div_smith = 0, mul_float = 0, mul_real = 0, sub_real = 0, abs_hypot = 0, norm_ok = 0;
    for (int i = 0; i < n; ++i) {
        volatile float va = rnd(&s);
        float a = va, b = rnd(&s), c = rnd(&s), d = rnd(&s), f = rnd(&s);
        std::complex<float> x(a, b), y(c, d);

        std::complex<float> q = x / y;
        double den = (double)c * c + (double)d * d;
        if (q.real() == (float)(((double)a * c + (double)b * d) / den) &&
            q.imag() == (float)(((double)b * c - (double)a * d) / den))
            ++div_double;
        float r1, i1;
        if (std::fabs(c) < std::fabs(d)) {
            float ratio = c / d, dn = (c * ratio) + d;
            r1 = ((a * ratio) + b) / dn;
            i1 = ((b * ratio) - a) / dn;
        } else {
            float ratio = d / c, dn = (d * ratio) + c;
            r1 = ((b * ratio) + a) / dn;
            i1 = (b - (a * ratio)) / dn;
        }
        if (q.real() == r1 && q.imag() == i1) ++div_smith;

        std::complex<float> m = x * y;
        if (m.real() == a * c - b * d && m.imag() == a * d + b * c) ++mul_float;
        std::complex<float> mr = x * f;
        if (mr.real() == a * f && mr.imag() == b * f) ++mul_real;
        std::complex<float> sr = f - x;
        if (sr.real() == f - a && sr.imag() == -b) ++sub_real;
        if (std::abs(x) == std::hypot(a, b)) ++abs_hypot;
        if (std::norm(x) == a * a + b * b) ++norm_ok;
    }
    std::printf("n %d\ndiv_double %d\ndiv_smith %d\nmul_float %d\nmul_real %d\nsub_real %d\nabs_hypot %d\nnorm %d\n", n,
                div_double, div_smith, mul_float, mul_real, sub_real, abs_hypot, norm_ok);
    return 0;
}
