// H0^(1)(x), H1^(1)(x) of real x > 0 in FP64, for the device and the host.
//
// Replaces the reference's BesselH0/BesselH1(double) = GSL J + iY (lib/bem2d_mathroutines.cpp:20-42).
// Three regimes, coefficients in hankel_tables.h (written by gen_hankel_tables.py from mpmath):
//   x < 2    ascending series in t = x^2/4 (A&S 9.1.10 and 9.1.11, n = 0, 1);
//   x >= 2   H_nu = sqrt(2/(pi x)) (P_nu + i Q_nu) exp(i(x - (2nu+1) pi/4)) (A&S 9.2.5-9.2.6),
//            P_nu and x Q_nu as degree-12 polynomials of z(y), y = (a/x)^2, on the pieces
//            [2, 2.6), [2.6, 3.5), [3.5, 5), [5, 8) and [8, inf) -- the last one is the
//            asymptotic regime, fitted over the whole half-line.
// The piece is selected by a switch so that every table index is a compile-time constant (the
// coefficients become instruction literals instead of a per-lane table).  sin/cos of the phase use
// sincos_reduced below.  Accuracy: about 3e-16 relative to |H| (pinned against mpmath by the tests).
#pragma once

#include <hip/hip_runtime.h>

#include "hankel_tables.h"

namespace bem2d_dev {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoOverPi = 0.63661977236758134308;
constexpr double kInvPi = 0.31830988618379067154;

template <int N>
__host__ __device__ __forceinline__ double horner(const double* c, double z)
{
    double r = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) r = fma(r, z, c[i]);
    return r;
}

// sin and cos of x for |x| < 2^30.  Cody-Waite reduction by pi/2 split into three doubles
// (pi/2 = P1 + P2 + P3 to 160 bits), then the Taylor series of sin and cos on |r| <= pi/4,
// truncated at r^17 and r^18 (remainder below 1e-19).  Both results within about 1 ulp.
__host__ __device__ __forceinline__ void sincos_reduced(double x, double& s, double& c)
{
    constexpr double P1 = 1.5707963267948966;
    constexpr double P2 = 6.123233995736766e-17;
    constexpr double P3 = -1.4973849048591698e-33;
    const double n = rint(x * 0.6366197723675814);
    double r = fma(-n, P1, x);
    r = fma(-n, P2, r);
    r = fma(-n, P3, r);
    const double r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - ... + r^16/17!)
    double ps = 2.8114572543455206e-15;           //  1/17!
    ps = fma(ps, r2, -7.6471637318198164e-13);    // -1/15!
    ps = fma(ps, r2, 1.6059043836821613e-10);     //  1/13!
    ps = fma(ps, r2, -2.5052108385441720e-08);    // -1/11!
    ps = fma(ps, r2, 2.7557319223985893e-06);     //  1/9!
    ps = fma(ps, r2, -1.9841269841269841e-04);    // -1/7!
    ps = fma(ps, r2, 8.3333333333333333e-03);     //  1/5!
    ps = fma(ps, r2, -1.6666666666666667e-01);    // -1/3!
    const double sr = fma(ps * r2, r, r);
    // cos r = 1 - r^2/2! + ... + r^18/18!
    double pc = -1.5619206968586225e-16;          // -1/18!
    pc = fma(pc, r2, 4.7794773323873853e-14);     //  1/16!
    pc = fma(pc, r2, -1.1470745597729725e-11);    // -1/14!
    pc = fma(pc, r2, 2.0876756987868099e-09);     //  1/12!
    pc = fma(pc, r2, -2.7557319223985888e-07);    // -1/10!
    pc = fma(pc, r2, 2.4801587301587302e-05);     //  1/8!
    pc = fma(pc, r2, -1.3888888888888889e-03);    // -1/6!
    pc = fma(pc, r2, 4.1666666666666667e-02);     //  1/4!
    pc = fma(pc, r2, -0.5);                       // -1/2!
    const double cr = fma(pc, r2, 1.0);
    const int q = ((int)(long long)n) & 3;
    if (q == 0) { s = sr; c = cr; }
    else if (q == 1) { s = cr; c = -sr; }
    else if (q == 2) { s = -sr; c = -cr; }
    else { s = -cr; c = sr; }
}

// Series regime, x < 2.
__host__ __device__ __forceinline__ void hankel01_series(double x, double& h0r, double& h0i,
                                                         double& h1r, double& h1i)
{
    const double t = 0.25 * x * x;
    const double hx = 0.5 * x;
    const double lg = log(hx);
    const double j0 = horner<BEM2D_HANKEL_NSER>(BEM2D_SER_J0, t);
    const double j1 = hx * horner<BEM2D_HANKEL_NSER>(BEM2D_SER_J1, t);
    const double y0 = kTwoOverPi * lg * j0 + horner<BEM2D_HANKEL_NSER>(BEM2D_SER_Y0, t);
    const double y1 = kTwoOverPi * lg * j1 - kTwoOverPi / x +
                      hx * horner<BEM2D_HANKEL_NSER>(BEM2D_SER_Y1, t);
    h0r = j0; h0i = y0; h1r = j1; h1i = y1;
}

// P_nu, Q_nu of piece P at x (invx = 1/x).
template <int P>
__host__ __device__ __forceinline__ void hankel_pq(double invx, double& P0, double& Q0, double& P1,
                                                   double& Q1)
{
    const double y = BEM2D_PIECE_A2[P] * invx * invx;
    const double z = (y - BEM2D_PIECE_YM[P]) * BEM2D_PIECE_YHINV[P];
    P0 = horner<BEM2D_HANKEL_DEG + 1>(BEM2D_FIT_P0[P], z);
    Q0 = horner<BEM2D_HANKEL_DEG + 1>(BEM2D_FIT_Q0[P], z) * invx;
    P1 = horner<BEM2D_HANKEL_DEG + 1>(BEM2D_FIT_P1[P], z);
    Q1 = horner<BEM2D_HANKEL_DEG + 1>(BEM2D_FIT_Q1[P], z) * invx;
}

// Modulus/phase combination for x >= 2.
__host__ __device__ __forceinline__ void hankel01_combine(double x, double invx, double P0,
                                                          double Q0, double P1, double Q1,
                                                          double& h0r, double& h0i, double& h1r,
                                                          double& h1i)
{
    double s, c;
    sincos_reduced(x, s, c);
    const double A = sqrt(kInvPi * invx); // sqrt(2/(pi x)) / sqrt(2)
    const double cp = c + s;             // sqrt(2) cos(x - pi/4)
    const double cm = s - c;             // sqrt(2) sin(x - pi/4)
    h0r = A * (P0 * cp - Q0 * cm);
    h0i = A * (P0 * cm + Q0 * cp);
    h1r = A * (P1 * cm + Q1 * cp);
    h1i = A * (Q1 * cm - P1 * cp);
}

__host__ __device__ __forceinline__ void hankel01(double x, double& h0r, double& h0i, double& h1r,
                                                  double& h1i)
{
    if (x < 2.0) {
        hankel01_series(x, h0r, h0i, h1r, h1i);
        return;
    }
    const double invx = 1.0 / x;
    double P0, Q0, P1, Q1;
    if (x >= 8.0)
        hankel_pq<4>(invx, P0, Q0, P1, Q1);
    else if (x >= 5.0)
        hankel_pq<3>(invx, P0, Q0, P1, Q1);
    else if (x >= 3.5)
        hankel_pq<2>(invx, P0, Q0, P1, Q1);
    else if (x >= 2.6)
        hankel_pq<1>(invx, P0, Q0, P1, Q1);
    else
        hankel_pq<0>(invx, P0, Q0, P1, Q1);
    hankel01_combine(x, invx, P0, Q0, P1, Q1, h0r, h0i, h1r, h1i);
}

} // namespace bem2d_dev
