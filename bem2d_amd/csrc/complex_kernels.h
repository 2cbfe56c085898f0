// Kernel functors for complex wavenumbers k = kr + i ki: the kernels' `k_.im != 0` branches
// (lib/bem2d_kernel.cpp:15-112) with Hankel values of complex argument from hankel01_complex.
#pragma once

#include "geometry.h"
#include "hankel_complex.h"

namespace bem2d_dev {

// Kernel functors for complex k = (kr, ki); tab / nt = contour table of arg k.
template <int KIND>
__host__ __device__ __forceinline__ void kernel_eval_ck(double kr, double ki, double eta,
                                                        const double* tab, int nt, double xx,
                                                        double xy, double yx, double yy,
                                                        double n1x, double n1y, double n2x,
                                                        double n2y, double& vr, double& vi)
{
    double wx, wy, d2, nw;
    {
#pragma clang fp contract(off)
        wx = xx - yx; // x - y
        wy = xy - yy;
        d2 = wx * wx + wy * wy;
        if (KIND == BEM2D_KERNEL_DL)
            nw = n2x * wx + n2y * wy;
        else
            nw = n1x * wx + n1y * wy;
    }
    const double d = sqrt(d2);
    double h0r, h0i, h1r, h1i;
    hankel01_complex(kr * d, ki * d, tab, nt, h0r, h0i, h1r, h1i);
    if (KIND == BEM2D_KERNEL_SL) { // (i/4) H0(kd)
        vr = -0.25 * h0i;
        vi = 0.25 * h0r;
        return;
    }
    // -(i k/4) H1(kd) c, c = n_y.(y - x)/d for DL, n_x.(x - y)/d otherwise
    const double c = (KIND == BEM2D_KERNEL_DL ? -nw : nw) / d;
    const double ar = kr * h1r - ki * h1i, ai = kr * h1i + ki * h1r; // k H1
    vr = 0.25 * c * ai;
    vi = -0.25 * c * ar;
    if (KIND == BEM2D_KERNEL_COMBINED) { // CDL - i eta (i/4) H0
        vr += 0.25 * eta * h0r;
        vi += 0.25 * eta * h0i;
    }
}

struct ComplexKernelEval {
    double kr, ki, eta;
    const double* tab;
    int nt;
    template <int KIND>
    __device__ __forceinline__ void eval(double xx, double xy, double yx, double yy, double n1x,
                                         double n1y, double n2x, double n2y, double& vr,
                                         double& vi) const
    {
        kernel_eval_ck<KIND>(kr, ki, eta, tab, nt, xx, xy, yx, yy, n1x, n1y, n2x, n2y, vr, vi);
    }
};

} // namespace bem2d_dev
