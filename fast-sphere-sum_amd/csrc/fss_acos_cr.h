// fss_acos_cr.h -- correctly rounded arccosine for the tree's near-tie MAC
// decisions, usable from host and device code.
//
// The reference decides the multipole acceptance criterion with the
// platform's acos (great_circ_dist, Old-BVE-code/helpers.h:211-215).  Two
// faithful libms can differ in the last bit, which flips a decision whose
// separation lies within a few ulp of theta.  Deciding those pairs with the
// correctly rounded value -- unique, so every correctly rounded libm agrees
// with it -- makes the interaction list independent of the libm.
//
// Method: a libm value y (within a few ulp) is moved to the neighbouring
// double while a rounding midpoint m between y and that neighbour lies on
// the wrong side of acos(x).  acos is decreasing, so acos(x) > m exactly when
// x < cos(m); cos(m) is evaluated in double-double arithmetic (Dekker, "A
// floating-point technique for extending the available precision", Numer.
// Math. 18, 1971), error about 1e-32, far below the distance between x and
// cos(m) for any midpoint m (cos of a nonzero rational is irrational, so it
// is never a double).  This is Ziv's rounding test (ACM TOMS 17, 1991) with
// the inverse function as the accurate evaluation.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define FSS_CR_HD __host__ __device__ inline
#else
#define FSS_CR_HD static inline
#endif

namespace fsscr {

struct dd {
    double hi, lo;
};

// a + b exactly as hi + lo (Knuth's two-sum)
FSS_CR_HD dd two_sum(double a, double b)
{
    double s = a + b;
    double v = s - a;
    return dd{s, (a - (s - v)) + (b - v)};
}

// a + b exactly, |a| >= |b|
FSS_CR_HD dd fast_two_sum(double a, double b)
{
    double s = a + b;
    return dd{s, b - (s - a)};
}

FSS_CR_HD dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

FSS_CR_HD dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }

FSS_CR_HD dd dd_mul(dd a, dd b)
{
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p); // the product's rounding error, exactly
    e += a.hi * b.lo;
    e += a.lo * b.hi;
    return fast_two_sum(p, e);
}

// a / q for a small positive integer q
FSS_CR_HD dd dd_div(dd a, double q)
{
    double h = a.hi / q;
    double r = fma(-h, q, a.hi) + a.lo;
    return fast_two_sum(h, r / q);
}

// 1 - r2 t / q
FSS_CR_HD dd horner_step(dd r2, dd t, double q)
{
    return dd_add(dd{1.0, 0.0}, dd_neg(dd_div(dd_mul(r2, t), q)));
}

// cos and sin of |r| <= 0.8 by their Taylor series in Horner form up to the
// terms of order 30 and 31 (the first omitted term is below 1e-38)
FSS_CR_HD dd dd_cos_small(dd r)
{
    dd r2 = dd_mul(r, r), t{1.0, 0.0};
    for (int n = 15; n >= 1; n--) t = horner_step(r2, t, (double)((2 * n - 1) * (2 * n)));
    return t;
}

FSS_CR_HD dd dd_sin_small(dd r)
{
    dd r2 = dd_mul(r, r), t{1.0, 0.0};
    for (int n = 15; n >= 1; n--) t = horner_step(r2, t, (double)((2 * n) * (2 * n + 1)));
    return dd_mul(r, t);
}

// cos(m) for m in [0, pi + 1e-15], m = k pi/2 + r with pi/2 to 161 bits
FSS_CR_HD dd dd_cos(dd m)
{
    const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
    const int k = (int)nearbyint(m.hi / P1);
    dd r = dd_add(m, dd{-k * P1, 0.0});
    r = dd_add(r, dd{-k * P2, 0.0});
    r = dd_add(r, dd{-k * P3, 0.0});
    if (k == 0) return dd_cos_small(r);
    if (k == 1) return dd_neg(dd_sin_small(r));
    return dd_neg(dd_cos_small(r));
}

// acos(x) correctly rounded to nearest, x in [-1, 1]
FSS_CR_HD double acos_cr(double x)
{
    if (x >= 1.0) return 0.0;
    if (x <= -1.0) return 3.141592653589793; // pi to nearest
    double y = acos(x);
    for (int it = 0; it < 8; it++) {
        const double up = nextafter(y, 4.0), dn = nextafter(y, -1.0);
        const dd cu = dd_cos(dd{y, 0.5 * (up - y)}); // midpoint to the next double up
        if (x - cu.hi < cu.lo) {                     // x < cos(mid): acos(x) above mid
            y = up;
            continue;
        }
        const dd cd = dd_cos(dd{y, 0.5 * (dn - y)}); // midpoint to the next double down
        if (x - cd.hi > cd.lo) {                     // x > cos(mid): acos(x) below mid
            y = dn;
            continue;
        }
        break;
    }
    return y;
}

} // namespace fsscr
