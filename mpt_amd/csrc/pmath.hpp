// Project FP64 sine/cosine/arccosine, identical on host and device.
//
// Method: the fdlibm algorithms (Sun Microsystems, 1993; freely
// redistributable), i.e. Cody-Waite reduction by pi/2 with a three-stage
// 33-bit split of pi/2 for |x| < 2^19*(pi/2), the minimax kernels
// __kernel_sin / __kernel_cos on [-pi/4, pi/4], and the rational
// approximation of e_acos.c.  Constants are fdlibm's published ones.
// Every operation is a plain IEEE-754 binary64 op and the code is compiled
// without FP contraction, so a host build and a gfx950 build return the
// same bits.  Arguments beyond the medium-range reduction (|x| >= 2^19*pi/2)
// are not needed by any scene in this project and return NaN.
#pragma once
#include "hd.hpp"

namespace mptg::pm {

MPTG_HD uint32_t high_word(double x) { return (uint32_t)(bits_of(x) >> 32); }

MPTG_HD double kernel_sin(double x, double y, int iy) {
    const double S1 = -1.66666666666666324348e-01;
    const double S2 = 8.33333333332248946124e-03;
    const double S3 = -1.98412698298579493134e-04;
    const double S4 = 2.75573137070700676789e-06;
    const double S5 = -2.50507602534068634195e-08;
    const double S6 = 1.58969099521155010221e-10;
    double z = x * x;
    double w = z * z;
    double r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6);
    double v = z * x;
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

MPTG_HD double kernel_cos(double x, double y) {
    const double C1 = 4.16666666666666019037e-02;
    const double C2 = -1.38888888888741095749e-03;
    const double C3 = 2.48015872894767294178e-05;
    const double C4 = -2.75573143513906633035e-07;
    const double C5 = 2.08757232129817482790e-09;
    const double C6 = -1.13596475577881948265e-11;
    double z = x * x;
    double w = z * z;
    double r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6));
    double hz = 0.5 * z;
    w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * r - x * y));
}

// Medium-range argument reduction: x = n*(pi/2) + (y0 + y1), |y0 + y1| <= pi/4
// (slightly more near the boundary).  Returns n mod 4 in [0,4), or -1 when |x|
// is outside the supported range (or not finite).
MPTG_HD int rem_pio2(double x, double* y0, double* y1) {
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    const double pio2_2 = 6.07710050630396597660e-11;
    const double pio2_2t = 2.02226624879595063154e-21;
    const double pio2_3 = 2.02226624871116645580e-21;
    const double pio2_3t = 8.47842766036889956997e-32;
    uint32_t hx = high_word(x);
    uint32_t ix = hx & 0x7fffffffu;
    if (ix >= 0x413921fbu) {  // |x| >= 2^19 * (pi/2), inf or nan
        *y0 = 0.0;
        *y1 = 0.0;
        return -1;
    }
    double t = x < 0 ? -x : x;
    int n = (int)(t * invpio2 + 0.5);
    double fn = (double)n;
    double r = t - fn * pio2_1;
    double w = fn * pio2_1t;
    int j = (int)(ix >> 20);
    double yy = r - w;
    int i = j - (int)((high_word(yy) >> 20) & 0x7ffu);
    if (i > 16) {  // cancellation: second stage
        t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        yy = r - w;
        i = j - (int)((high_word(yy) >> 20) & 0x7ffu);
        if (i > 49) {  // third stage
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            yy = r - w;
        }
    }
    double yl = (r - yy) - w;
    if (x < 0) {
        *y0 = -yy;
        *y1 = -yl;
        return (-n) & 3;
    }
    *y0 = yy;
    *y1 = yl;
    return n & 3;
}

// sin and cos of one argument.
MPTG_HD void sincos(double x, double* s, double* c) {
    uint32_t ix = high_word(x) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {  // |x| <~ pi/4
        if (ix < 0x3e400000u) {  // |x| < 2^-27
            *s = x;
            *c = 1.0;
            return;
        }
        *s = kernel_sin(x, 0.0, 0);
        *c = kernel_cos(x, 0.0);
        return;
    }
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    if (n < 0) {
        double nan = double_of(0x7ff8000000000000ull);
        *s = nan;
        *c = nan;
        return;
    }
    double ks = kernel_sin(y0, y1, 1);
    double kc = kernel_cos(y0, y1);
    switch (n) {
        case 0: *s = ks; *c = kc; break;
        case 1: *s = kc; *c = -ks; break;
        case 2: *s = -ks; *c = -kc; break;
        default: *s = -kc; *c = ks; break;
    }
}

MPTG_HD double sin(double x) {
    double s, c;
    sincos(x, &s, &c);
    return s;
}

MPTG_HD double cos(double x) {
    double s, c;
    sincos(x, &s, &c);
    return c;
}

// arccos (fdlibm e_acos.c): rational approximation R(z) = p(z)/q(z) of
// (asin(sqrt z) - sqrt z)/(z sqrt z) combined with sqrt.
MPTG_HD double acos(double x) {
    const double pi = 3.14159265358979311600e+00;
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    const double pS0 = 1.66666666666666657415e-01;
    const double pS1 = -3.25565818622400915405e-01;
    const double pS2 = 2.01212532134862925881e-01;
    const double pS3 = -4.00555345006794114027e-02;
    const double pS4 = 7.91534994289814532176e-04;
    const double pS5 = 3.47933107596021167570e-05;
    const double qS1 = -2.40339491173441421878e+00;
    const double qS2 = 2.02094576023350569471e+00;
    const double qS3 = -6.88283971605453293030e-01;
    const double qS4 = 7.70381505559019352791e-02;
    uint64_t bx = bits_of(x);
    uint32_t hx = (uint32_t)(bx >> 32);
    uint32_t ix = hx & 0x7fffffffu;
    if (ix >= 0x3ff00000u) {  // |x| >= 1
        if (((ix - 0x3ff00000u) | (uint32_t)bx) == 0) {
            if ((int32_t)hx > 0) return 0.0;
            return pi + 2.0 * pio2_lo;
        }
        return double_of(0x7ff8000000000000ull);
    }
    if (ix < 0x3fe00000u) {  // |x| < 0.5
        if (ix <= 0x3c600000u) return pio2_hi + pio2_lo;  // |x| < 2^-57
        double z = x * x;
        double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        double r = p / q;
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if ((int32_t)hx < 0) {  // x < -0.5
        double z = (1.0 + x) * 0.5;
        double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        double s = __builtin_sqrt(z);
        double r = p / q;
        double w = r * s - pio2_lo;
        return pi - 2.0 * (s + w);
    }
    // x > 0.5
    double z = (1.0 - x) * 0.5;
    double s = __builtin_sqrt(z);
    double df = double_of(bits_of(s) & 0xffffffff00000000ull);
    double c = (z - df * df) / (s + df);
    double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    double r = p / q;
    double w = r * s + c;
    return 2.0 * (df + w);
}

}  // namespace mptg::pm
