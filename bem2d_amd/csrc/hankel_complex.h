// H0^(1)(z) and H1^(1)(z) of complex z with Re z > 0, in FP64, identical code on the device and the
// host.  Stands in for AMOS zbesh (kode=1, n=1) behind BesselH0(complex) / BesselH1(complex)
// (lib/bem2d_mathroutines.cpp:44-75), which the kernels call when the wavenumber has a nonzero
// imaginary part (lib/bem2d_kernel.cpp:15-112).  Three regions:
//
//   |z| < 2        ascending series (A&S 9.1.10, 9.1.11 with n = 0, 1), 16 terms in t = -z^2/4;
//   2 <= |z| < 20  the Sommerfeld integral H_nu(z) = (1/(pi i)) int_{-inf}^{inf+pi i}
//                  exp(z sinh w - nu w) dw (A&S 9.1.25 form; DLMF 10.9.18) on the path
//                  w(t) = t + i(pi/2 + c gd(t)), gd = Gudermannian, c = 1 - 2 arg(z)/pi, by the
//                  trapezoidal rule in t.  The path leaves the saddle w = i pi/2 and ends in the
//                  valleys of exp(z sinh w) at both ends for every Re z > 0, so the integrand decays
//                  doubly exponentially and the rule converges geometrically in 1/h.  With
//                  S(t) = sinh w(t) - i the sum is H_nu = e^{iz} sum_t Q_nu(t) exp(z S(t)), and the
//                  node table (S, Q_0, Q_1) depends only on arg z -- i.e. on arg k within one
//                  assembly -- so it is built once per wavenumber on the host (contour_table);
//   |z| >= 20      Hankel's asymptotic expansion (A&S 9.2.5-9.2.10), summed until the terms
//                  fall below 1e-17 of the sum (at most 30).
//
// Accuracy (tests, against mpmath at 30 digits and scipy's AMOS): a few 1e-15 relative to |H| for
// |arg z| <= 0.95 pi/2 over 1e-6 <= |z| <= 1e4 wherever H is a finite double.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

namespace bem2d_dev {

constexpr int kContourMaxNodes = 192;
constexpr int kContourDoubles = 6; // per node: S (re, im), Q0 (re, im), Q1 (re, im)

// Trapezoidal nodes of the contour for arg z = phi (|phi| < pi/2): h = 0.1 when phi >= 0, 0.05
// below (the path then turns against the oscillation of the integrand), t in [-4.5, 4.5].
// Writes kContourDoubles * count doubles to tab and returns count (<= kContourMaxNodes).
inline int contour_table(double phi, double* tab)
{
    const double PI = 3.14159265358979323846;
    const double h = phi >= 0.0 ? 0.1 : 0.05, T = 4.5;
    const double c = 1.0 - 2.0 * phi / PI;
    const int count = (int)lround(2.0 * T / h) + 1;
    for (int m = 0; m < count; ++m) {
        const double t = -T + m * h;
        const double th = PI / 2 + c * atan(sinh(t)); // Im w
        const double sh = sinh(t), ch = cosh(t);
        // sinh(t + i th) - i
        const double sr = sh * cos(th), si = ch * sin(th) - 1.0;
        // w'(t) = 1 + i c sech t;  Q0 = h w' / (pi i) = (h/pi)(c sech t - i)
        const double dwr = 1.0, dwi = c / ch;
        const double q0r = h / PI * dwi, q0i = -h / PI * dwr;
        // Q1 = Q0 e^{-w},  e^{-w} = e^{-t} (cos th - i sin th)
        const double er = exp(-t) * cos(th), ei = -exp(-t) * sin(th);
        double* o = tab + kContourDoubles * m;
        o[0] = sr;
        o[1] = si;
        o[2] = q0r;
        o[3] = q0i;
        o[4] = q0r * er - q0i * ei;
        o[5] = q0r * ei + q0i * er;
    }
    return count;
}

__host__ __device__ __forceinline__ void cplx_sincos(double a, double& s, double& c)
{
#ifdef __HIP_DEVICE_COMPILE__
    sincos(a, &s, &c);
#else
    s = sin(a);
    c = cos(a);
#endif
}

// (h0r, h0i) = H0^(1)(z), (h1r, h1i) = H1^(1)(z) for z = zr + i zi, zr > 0.  tab / nt: the contour
// table of arg z (contour_table), read only for 2 <= |z| < 20.
__host__ __device__ inline void hankel01_complex(double zr, double zi, const double* tab, int nt,
                                                 double& h0r, double& h0i, double& h1r,
                                                 double& h1i)
{
    const double PI = 3.14159265358979323846;
    const double r2 = zr * zr + zi * zi;
    if (r2 < 4.0) {
        // t = -z^2/4; term_m = t^m/(m!)^2; Hm = 1 + 1/2 + ... + 1/m
        const double tr = -0.25 * (zr * zr - zi * zi), ti = -0.5 * zr * zi;
        double ur = 1.0, ui = 0.0, hm = 0.0;
        double j0r = 1.0, j0i = 0.0, s0r = 0.0, s0i = 0.0;
        double j1r = 1.0, j1i = 0.0, s1r = 1.0, s1i = 0.0; // m = 0: term/(m+1), (2Hm + 1/(m+1))...
        for (int m = 1; m < 16; ++m) {
            const double inv = 1.0 / ((double)m * (double)m);
            const double nr = (ur * tr - ui * ti) * inv, ni = (ur * ti + ui * tr) * inv;
            ur = nr;
            ui = ni;
            hm += 1.0 / m;
            const double f1 = 1.0 / (m + 1);
            j0r += ur;
            j0i += ui;
            s0r += hm * ur;
            s0i += hm * ui;
            j1r += ur * f1;
            j1i += ui * f1;
            const double g = (2.0 * hm + f1) * f1;
            s1r += g * ur;
            s1i += g * ui;
        }
        // J1 = (z/2) sum term/(m+1);  S1' = (z/2) sum (2Hm + 1/(m+1)) term/(m+1)
        const double hzr = 0.5 * zr, hzi = 0.5 * zi;
        const double J1r = hzr * j1r - hzi * j1i, J1i = hzr * j1i + hzi * j1r;
        const double T1r = hzr * s1r - hzi * s1i, T1i = hzr * s1i + hzi * s1r;
        // L = ln(z/2) + gamma
        const double Lr = 0.5 * log(0.25 * r2) + 0.57721566490153286061, Li = atan2(zi, zr);
        // Y0 = (2/pi)(L J0 - S0)
        const double y0r = 2.0 / PI * (Lr * j0r - Li * j0i - s0r);
        const double y0i = 2.0 / PI * (Lr * j0i + Li * j0r - s0i);
        // Y1 = (2/pi) L J1 - 2/(pi z) - (1/pi) S1'
        const double izr = zr / r2, izi = -zi / r2; // 1/z
        const double y1r = 2.0 / PI * (Lr * J1r - Li * J1i) - 2.0 / PI * izr - T1r / PI;
        const double y1i = 2.0 / PI * (Lr * J1i + Li * J1r) - 2.0 / PI * izi - T1i / PI;
        // H = J + iY
        h0r = j0r - y0i;
        h0i = j0i + y0r;
        h1r = J1r - y1i;
        h1i = J1i + y1r;
        return;
    }
    double ez, cz, sz; // e^{iz} = ez (cz + i sz)
    ez = exp(-zi);
    cplx_sincos(zr, sz, cz);
    if (r2 < 400.0) {
        double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
        for (int m = 0; m < nt; ++m) {
            const double* o = tab + kContourDoubles * m;
            const double er = zr * o[0] - zi * o[1];
            if (er < -45.0) continue; // below 1e-19 of the saddle's term
            const double ei = zr * o[1] + zi * o[0];
            double s, c;
            cplx_sincos(ei, s, c);
            const double mag = exp(er);
            const double vr = mag * c, vi = mag * s;
            a0r += o[2] * vr - o[3] * vi;
            a0i += o[2] * vi + o[3] * vr;
            a1r += o[4] * vr - o[5] * vi;
            a1i += o[4] * vi + o[5] * vr;
        }
        const double pr = ez * cz, pi_ = ez * sz;
        h0r = pr * a0r - pi_ * a0i;
        h0i = pr * a0i + pi_ * a0r;
        h1r = pr * a1r - pi_ * a1i;
        h1i = pr * a1i + pi_ * a1r;
        return;
    }
    // Hankel expansion: H_nu = sqrt(2/(pi z)) e^{i(z - nu pi/2 - pi/4)} sum_k i^k a_k(nu) z^-k,
    // term_k = term_{k-1} * i (4 nu^2 - (2k-1)^2) / (8 k z)
    const double izr = zr / r2, izi = -zi / r2;
    double p0r = 1.0, p0i = 0.0, p1r = 1.0, p1i = 0.0;
    double t0r = 1.0, t0i = 0.0, t1r = 1.0, t1i = 0.0;
    for (int kk = 1; kk <= 30; ++kk) {
        const double o2 = (2.0 * kk - 1.0) * (2.0 * kk - 1.0);
        const double f0 = -o2 / (8.0 * kk), f1 = (4.0 - o2) / (8.0 * kk);
        // multiply by i f / z:  i/z = (-izi, izr)
        const double mr = -izi, mi = izr;
        double nr = (t0r * mr - t0i * mi) * f0, ni = (t0r * mi + t0i * mr) * f0;
        t0r = nr;
        t0i = ni;
        nr = (t1r * mr - t1i * mi) * f1;
        ni = (t1r * mi + t1i * mr) * f1;
        t1r = nr;
        t1i = ni;
        p0r += t0r;
        p0i += t0i;
        p1r += t1r;
        p1i += t1i;
        const double tt = t0r * t0r + t0i * t0i + t1r * t1r + t1i * t1i;
        if (tt < 1e-34 * (p0r * p0r + p0i * p0i)) break;
    }
    // sqrt(2/(pi z)) = sqrt(2/pi) / sqrt(z);  sqrt(z) = a + i zi/(2a), a = sqrt((|z| + zr)/2)
    const double a = sqrt(0.5 * (sqrt(r2) + zr)), b = zi / (2.0 * a);
    const double sq = sqrt(2.0 / PI) / (a * a + b * b);
    const double qr = sq * a, qi = -sq * b; // sqrt(2/pi) conj(sqrt z)/|sqrt z|^2
    // e^{i(z - pi/4)} = e^{iz} (1 - i)/sqrt(2)
    const double R2 = 0.70710678118654752440;
    const double er = ez * (cz + sz) * R2, ei = ez * (sz - cz) * R2;
    const double fr = qr * er - qi * ei, fi = qr * ei + qi * er;
    h0r = fr * p0r - fi * p0i;
    h0i = fr * p0i + fi * p0r;
    // H1 carries the extra factor e^{-i pi/2} = -i
    const double gr = fr * p1r - fi * p1i, gi = fr * p1i + fi * p1r;
    h1r = gi;
    h1i = -gr;
}

} // namespace bem2d_dev
