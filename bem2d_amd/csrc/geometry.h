// Element Map / Deriv / Normal on the device (and host), keyed by element type.
//
// Mirrors the reference's curve and element classes so that node positions agree with it to the
// last bits: the association of products such as 2*PI*t (evaluated as (2*PI)*t) is kept, and
// floating-point contraction is off (plain IEEE multiply and add, as the reference's x86-64 build
// evaluates them).  Near-coincident quadrature nodes make the integrand sensitive to the last bit
// of the node positions, so fused multiply-adds here would move the matrix by ~1e-11.
//   Circle, Trefoil, Kite, Ellipse, EllipseArc, InvEllipse   lib/bem2d_curvetypes.h:13-99
//   ConstElement, AnalyticCurveElement                       lib/bem2d_element.h:72-116
//   Element::Normal = normalize(D.y, -D.x)                    lib/bem2d_element.h:29-34
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bem2d_hip.h"
#include "hankel.h"

namespace bem2d_dev {

constexpr double kRefPi = 3.14159265358979323846; // lib/bem2d_defs.h: PI

struct Complex2 {
    double re, im;
};

// ---- correctly rounded sin and cos of curve angles ---------------------------------------------
// Near-coincident quadrature nodes make the double-layer factor n.(x - y) ~ kappa d^2 / 2 cancel,
// so that a one-ulp difference in a node position moves near-field entries by ~ulp / (kappa d^2)
// (1e-9 relative at ten elements per wavelength on N = 7000 elements).  The reference evaluates
// its curves with the C library's sin and cos, which round correctly in all but a tiny fraction of
// cases; so do these: the argument is reduced by pi/2 in double-double arithmetic (pi/2 to 159
// bits), the series of sin and cos are summed with their leading terms in double-double (error
// below 2^-70 relative), and the result is rounded once.  Contraction is off throughout: a product
// fused into a two-sum would break its error-free transformation.
struct DD {
    double hi, lo;
};
__host__ __device__ __forceinline__ DD dd_two_sum(double a, double b)
{
#pragma clang fp contract(off)
    const double s = a + b, bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ __forceinline__ DD dd_fast(double a, double b) // |a| >= |b|
{
#pragma clang fp contract(off)
    const double s = a + b;
    return DD{s, b - (s - a)};
}
__host__ __device__ __forceinline__ DD dd_prod(double a, double b)
{
#pragma clang fp contract(off)
    const double p = a * b;
    return DD{p, fma(a, b, -p)};
}
__host__ __device__ __forceinline__ DD dd_add(DD a, DD b)
{
#pragma clang fp contract(off)
    const DD s = dd_two_sum(a.hi, b.hi);
    return dd_fast(s.hi, s.lo + (a.lo + b.lo));
}
__host__ __device__ __forceinline__ DD dd_mul(DD a, DD b)
{
#pragma clang fp contract(off)
    const DD p = dd_prod(a.hi, b.hi);
    return dd_fast(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}

__host__ __device__ __forceinline__ void sincos_cr(double x, double& s, double& c)
{
#pragma clang fp contract(off)
    // pi/2 = P1 + P2 + P3, 1/k! as hi + lo (k <= 7) and as doubles beyond
    constexpr double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17,
                     P3 = -1.4973849048591698e-33, TWO_OVER_PI = 0.6366197723675814;
    const double n = rint(x * TWO_OVER_PI);
    // r = x - n (P1 + P2 + P3): the products exact as double-double pairs
    const DD a1 = dd_prod(n, P1), a2 = dd_prod(n, P2);
    DD r = dd_two_sum(x, -a1.hi);
    r = dd_add(r, DD{-a1.lo, 0.0});
    r = dd_add(r, DD{-a2.hi, -a2.lo});
    r = dd_add(r, DD{-n * P3, 0.0});
    const DD r2 = dd_mul(r, r);
    const double z = r2.hi;
    // sin r = r + r^3 (S3 + r^2 (S5 + r^2 (S7 + r^2 ts)))
    double ts;
    {
        // the double tail: S9 .. S29 (S29 first)
        const double t[11] = {1.1309962886447716e-31, -9.183689863795546e-29,
                              6.446950284384474e-26, -3.868170170630684e-23,
                              1.9572941063391263e-20, -8.22063524662433e-18,
                              2.8114572543455206e-15, -7.647163731819816e-13,
                              1.6059043836821613e-10, -2.505210838544172e-08,
                              2.7557319223985893e-06};
        ts = t[0];
#pragma unroll
        for (int i = 1; i < 11; ++i) ts = fma(ts, z, t[i]);
    }
    DD us = dd_add(DD{-0.0001984126984126984, -1.7209558293420705e-22}, DD{ts * z, 0.0});
    us = dd_add(DD{0.008333333333333333, 1.1564823173178714e-19}, dd_mul(r2, us));
    us = dd_add(DD{-0.16666666666666666, -9.25185853854297e-18}, dd_mul(r2, us));
    const DD sr = dd_add(r, dd_mul(dd_mul(r, r2), us));
    // cos r = 1 + r^2 (C2 + r^2 (C4 + r^2 (C6 + r^2 tc)))
    double tc;
    {
        const double t[11] = {3.279889237069838e-30, -2.4795962632247976e-27,
                              1.6117375710961184e-24, -8.896791392450574e-22,
                              4.110317623312165e-19, -1.5619206968586225e-16,
                              4.779477332387385e-14, -1.1470745597729725e-11,
                              2.08767569878681e-09, -2.755731922398589e-07,
                              2.48015873015873e-05};
        tc = t[0];
#pragma unroll
        for (int i = 1; i < 11; ++i) tc = fma(tc, z, t[i]);
    }
    DD uc = dd_add(DD{-0.001388888888888889, 5.300543954373577e-20}, DD{tc * z, 0.0});
    uc = dd_add(DD{0.041666666666666664, 2.3129646346357427e-18}, dd_mul(r2, uc));
    uc = dd_add(DD{-0.5, 0.0}, dd_mul(r2, uc));
    const DD cr = dd_add(DD{1.0, 0.0}, dd_mul(r2, uc));
    const double sv = sr.hi + sr.lo, cv = cr.hi + cr.lo;
    const int q = ((int)(long long)n) & 3;
    const double a = (q & 1) ? cv : sv; // sin before the sign
    const double b = (q & 1) ? sv : cv; // cos before the sign
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// sin and cos of a curve angle: correctly rounded on the device, the C library on the host.
__host__ __device__ __forceinline__ void angle_sincos(double a, double& s, double& c)
{
#ifdef __HIP_DEVICE_COMPILE__
    sincos_cr(a, s, c);
#else
    s = sin(a);
    c = cos(a);
#endif
}

// Curve Map/Deriv at curve parameter t; c = curve constructor arguments.
__host__ __device__ __forceinline__ void curve_eval(int type, const double* c, double t, double& mx,
                                                    double& my, double& dx, double& dy)
{
#pragma clang fp contract(off)
    const double PI = kRefPi;
    switch (type) {
    case BEM2D_ELEM_CIRCLE:
    {
        double s1, c1;
        angle_sincos(2 * PI * t, s1, c1);
        mx = c1;
        my = s1;
        dx = 2 * PI * (-1.0 * s1);
        dy = 2 * PI * c1;
        break;
    }
    case BEM2D_ELEM_TREFOIL: {
        double s1, c1, s3, c3;
        angle_sincos(2 * PI * t, s1, c1);
        angle_sincos(3 * 2 * PI * t, s3, c3);
        const double r = 1 + 0.3 * c3;
        mx = r * c1;
        my = r * s1;
        const double rp = -0.3 * 3 * 2 * PI * s3;
        dx = rp * c1 - r * 2 * PI * s1;
        dy = rp * s1 + r * 2 * PI * c1;
        break;
    }
    case BEM2D_ELEM_KITE: {
        double s1, c1, s2, c2;
        angle_sincos(2 * PI * t, s1, c1);
        angle_sincos(4 * PI * t, s2, c2);
        mx = c1 + 0.65 * c2 - 0.65;
        my = 1.5 * s1;
        dx = -2 * PI * s1 - 0.65 * 4 * PI * s2;
        dy = 1.5 * 2 * PI * c1;
        break;
    }
    case BEM2D_ELEM_ELLIPSE: {
        const double rho = c[0];
        double s1, c1;
        angle_sincos(2 * PI * t, s1, c1);
        mx = rho * c1;
        my = 1 / rho * s1;
        dx = -rho * 2 * PI * s1;
        dy = 2 * PI / rho * c1;
        break;
    }
    case BEM2D_ELEM_ELLIPSE_ARC: {
        const double a = c[0], b = c[1], t1 = c[2], t2 = c[3];
        double s1, c1;
        angle_sincos(t1 + t * (t2 - t1), s1, c1);
        mx = a * c1;
        my = b * s1;
        dx = -a * (t2 - t1) * s1;
        dy = b * (t2 - t1) * c1;
        break;
    }
    case BEM2D_ELEM_INV_ELLIPSE: {
        // res = e^s / (1 + alpha e^{2s}),  res' = 2 pi i e^s (1 - alpha e^{2s}) / (1 + alpha e^{2s})^2
        const double alpha = c[0];
        const double th = 2.0 * PI * t;
        double es_i, es_r, s2, c2;
        angle_sincos(th, es_i, es_r);
        angle_sincos(2.0 * th, s2, c2);
        const double e2r = alpha * c2, e2i = alpha * s2;
        const double den_r = 1.0 + e2r, den_i = e2i;
        const double dd = den_r * den_r + den_i * den_i;
        // e^s / den
        const double qr = (es_r * den_r + es_i * den_i) / dd;
        const double qi = (es_i * den_r - es_r * den_i) / dd;
        mx = qr;
        my = qi;
        // (2 pi i) * q / den * (1 - e2)
        const double q2r = (qr * den_r + qi * den_i) / dd;
        const double q2i = (qi * den_r - qr * den_i) / dd;
        const double fr = 1.0 - e2r, fi = -e2i;
        const double pr = q2r * fr - q2i * fi, pi_ = q2r * fi + q2i * fr;
        dx = -2.0 * PI * pi_;
        dy = 2.0 * PI * pr;
        break;
    }
    default:
        mx = my = dx = dy = __builtin_nan("");
    }
}

// Element Map/Deriv at local parameter t in [0, 1]; p = the element's 8 parameters.
__host__ __device__ __forceinline__ void elem_eval(int type, const double* p, double t, double& mx,
                                                   double& my, double& dx, double& dy)
{
#pragma clang fp contract(off)
    if (type == BEM2D_ELEM_FLAT) {
        mx = p[0] + t * p[2];
        my = p[1] + t * p[3];
        dx = p[2];
        dy = p[3];
        return;
    }
    const double ts = p[0], te = p[1];
    double cdx, cdy;
    curve_eval(type, p + 2, ts + t * (te - ts), mx, my, cdx, cdy);
    dx = (te - ts) * cdx;
    dy = (te - ts) * cdy;
}

// Unit normal (D.y, -D.x)/|D| and arc-length factor |D|.
__host__ __device__ __forceinline__ void normal_len(double dx, double dy, double& nx, double& ny,
                                                    double& len)
{
#pragma clang fp contract(off)
    len = sqrt(dx * dx + dy * dy); // equals length(Point(D.y, -D.x)) bit for bit
    nx = dy / len;
    ny = -dx / len;
}

} // namespace bem2d_dev
