/* The quotient steps of the FP64 sweeps' WENO5 form on the CPU (no GPU):
 * rcp_refined / quot_rcp / recon_weno5_rcp of ns3d_cpp_amd/csrc/ns3d_device.hpp
 * restated with C's fma(), against C's correctly rounded division and the
 * literal recon_weno5 (reconstruction.cpp:93-124).  The device seeds the
 * Newton steps with v_rcp_f64; here the seed is 1/b perturbed by a random
 * relative error of up to 2^-20 (the steps must not depend on its exact
 * bits).  Cases:
 *   1. random quotients a / b inside the range the device comment states
 *      (a above 2^-969, a / b normal, b below 2^1022, a / b below 2^768 b);
 *   2. WENO5 on random stencils whose values are 0 or within [2^-900, 2^200]
 *      in magnitude, both forms in the same expression order, compared bit
 *      for bit (a -0 against +0 counted apart);
 *   3. the same forms' 1/6 quotients of exact integer combinations.
 * Prints the counts and exits non-zero on any mismatch.  Built with
 * -ffp-contract=off like the FP64 device path; -DPLANT_NO_CORRECTION builds
 * the steps without their last correction, which must fail. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint64_t next(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
}
static double unif(void) { return (double)(next() >> 11) * 0x1p-53; }

static double seed_rcp(double b)
{
    return (1.0 / b) * (1.0 + (2.0 * unif() - 1.0) * 0x1p-20);
}

static double rcp_refined(double b)
{
    double r = seed_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return r;
}

static double quot_rcp(double a, double b, double r)
{
    const double q = a * r;
#ifdef PLANT_NO_CORRECTION   /* a planted mistake: the last correction left out */
    (void)b;
    return q;
#else
    return fma(fma(-b, q, a), r, q);
#endif
}

static double weno5_literal(const double *s)
{
    double f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
    double eps = 1e-6;
    double beta0 = 13.0 / 12.0 * (f0 - 2 * f1 + f2) * (f0 - 2 * f1 + f2)
                 + 0.25 * (f0 - 4 * f1 + 3 * f2) * (f0 - 4 * f1 + 3 * f2);
    double beta1 = 13.0 / 12.0 * (f1 - 2 * f2 + f3) * (f1 - 2 * f2 + f3)
                 + 0.25 * (f1 - f3) * (f1 - f3);
    double beta2 = 13.0 / 12.0 * (f2 - 2 * f3 + f4) * (f2 - 2 * f3 + f4)
                 + 0.25 * (3 * f2 - 4 * f3 + f4) * (3 * f2 - 4 * f3 + f4);
    double alpha0 = 0.1 / ((eps + beta0) * (eps + beta0));
    double alpha1 = 0.6 / ((eps + beta1) * (eps + beta1));
    double alpha2 = 0.3 / ((eps + beta2) * (eps + beta2));
    double asum = alpha0 + alpha1 + alpha2;
    double w0 = alpha0 / asum;
    double w1 = alpha1 / asum;
    double w2 = alpha2 / asum;
    double p0 = (2 * f0 - 7 * f1 + 11 * f2) / 6.0;
    double p1 = (-f1 + 5 * f2 + 2 * f3) / 6.0;
    double p2 = (2 * f2 + 5 * f3 - f4) / 6.0;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

static double weno5_rcp(const double *s, double r6)
{
    double f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
    double eps = 1e-6;
    double beta0 = 13.0 / 12.0 * (f0 - 2 * f1 + f2) * (f0 - 2 * f1 + f2)
                 + 0.25 * (f0 - 4 * f1 + 3 * f2) * (f0 - 4 * f1 + 3 * f2);
    double beta1 = 13.0 / 12.0 * (f1 - 2 * f2 + f3) * (f1 - 2 * f2 + f3)
                 + 0.25 * (f1 - f3) * (f1 - f3);
    double beta2 = 13.0 / 12.0 * (f2 - 2 * f3 + f4) * (f2 - 2 * f3 + f4)
                 + 0.25 * (3 * f2 - 4 * f3 + f4) * (3 * f2 - 4 * f3 + f4);
    double d0 = (eps + beta0) * (eps + beta0);
    double d1 = (eps + beta1) * (eps + beta1);
    double d2 = (eps + beta2) * (eps + beta2);
    double alpha0 = quot_rcp(0.1, d0, rcp_refined(d0));
    double alpha1 = quot_rcp(0.6, d1, rcp_refined(d1));
    double alpha2 = quot_rcp(0.3, d2, rcp_refined(d2));
    double asum = alpha0 + alpha1 + alpha2;
    const double ra = rcp_refined(asum);
    double w0 = quot_rcp(alpha0, asum, ra);
    double w1 = quot_rcp(alpha1, asum, ra);
    double w2 = quot_rcp(alpha2, asum, ra);
    double p0 = quot_rcp(2 * f0 - 7 * f1 + 11 * f2, 6.0, r6);
    double p1 = quot_rcp(-f1 + 5 * f2 + 2 * f3, 6.0, r6);
    double p2 = quot_rcp(2 * f2 + 5 * f3 - f4, 6.0, r6);
    return w0 * p0 + w1 * p1 + w2 * p2;
}

static int same_bits(double x, double y)
{
    uint64_t a, b;
    memcpy(&a, &x, 8);
    memcpy(&b, &y, 8);
    return a == b;
}

/* a value of magnitude 2^e (e uniform in [lo, hi]) times a random
   significand, random sign; zero with probability pz */
static double rand_val(int lo, int hi, double pz)
{
    if (unif() < pz) return unif() < 0.5 ? 0.0 : -0.0;
    const int e = lo + (int)(next() % (uint64_t)(hi - lo + 1));
    const double v = ldexp(1.0 + unif(), e);
    return unif() < 0.5 ? -v : v;
}

int main(void)
{
    long bad_q = 0, n_q = 0, bad_w = 0, zero_sign = 0, n_w = 0, bad_p = 0, n_p = 0;
    /* 1. generic quotients in the stated range */
    for (long t = 0; t < 8000000; ++t) {
        const double b = rand_val(-1000, 1000, 0.0);
        const int eq = -1000 + (int)(next() % 1700u);          /* a / b in [2^-1000, 2^700) */
        const double a = b * ldexp(1.0 + unif(), eq);
        if (!(fabs(a) >= 0x1p-968) || !isfinite(a) || fabs(a / b) < 0x1p-1021) continue;
        ++n_q;
        if (!same_bits(quot_rcp(a, b, rcp_refined(b)), a / b)) ++bad_q;
    }
    /* 2. WENO5 on stencils in the guard's range */
    const double r6 = rcp_refined(6.0);
    for (long t = 0; t < 3000000; ++t) {
        double s[5];
        const int mode = (int)(t % 3);
        const int base = -900 + (int)(next() % 1090u);       /* one magnitude, or spread */
        for (int q = 0; q < 5; ++q) {
            if (mode == 0) s[q] = rand_val(-900, 200, 0.1);
            else if (mode == 1) s[q] = rand_val(base, base + 10 < 200 ? base + 10 : 200, 0.05);
            else s[q] = (double)((int)(next() % 7u) - 3) * ldexp(1.0, base);   /* exact, ties and zeros */
        }
        ++n_w;
        const double x = weno5_literal(s), y = weno5_rcp(s, r6);
        if (!same_bits(x, y)) {
            if (x == y) ++zero_sign;
            else ++bad_w;
        }
    }
    /* 3. the 1/6 quotients of integer combinations near the range's ends */
    for (long t = 0; t < 3000000; ++t) {
        const double x = (double)((long)(next() % 2000001u) - 1000000) * ldexp(1.0, -952 + (int)(next() % 1150u));
        if (x == 0.0 || fabs(x) >= 0x1p205) continue;
        ++n_p;
        if (!same_bits(quot_rcp(x, 6.0, r6), x / 6.0)) ++bad_p;
    }
    printf("quotients %ld: %ld differ\n", n_q, bad_q);
    printf("weno5 stencils %ld: %ld differ, %ld differ only in a zero's sign\n", n_w, bad_w, zero_sign);
    printf("sixths %ld: %ld differ\n", n_p, bad_p);
    return (bad_q || bad_w || bad_p) ? 1 : 0;
}
