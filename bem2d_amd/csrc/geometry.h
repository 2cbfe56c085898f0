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

// sin and cos of a curve angle: the reduced-argument routine on the device (few registers), the C
// library on the host.
__host__ __device__ __forceinline__ void angle_sincos(double a, double& s, double& c)
{
#ifdef __HIP_DEVICE_COMPILE__
    sincos_reduced(a, s, c);
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
