// Drop-in replacement of lib/bem2d_curve.h:13-23.  Curve::Length is computed with composite
// 16-point Gauss-Legendre on 64 panels (the reference uses GSL's adaptive qag, epsrel 1e-4,
// lib/bem2d_curve.cpp:15-28).
//
// device_desc() is new: a curve that the GPU can evaluate by itself reports its element-type tag
// and constructor arguments (include/bem2d_hip.h); user-defined curves return false and are
// rejected by the device assembly.
#ifndef BEM2D_DROPIN_CURVE_H
#define BEM2D_DROPIN_CURVE_H

#include "bem2d_defs.h"
#include "bem2d_point.h"

namespace bem2d {

class Curve {
public:
    virtual Point Map(double t) const = 0;
    virtual Point Deriv(double t) const = 0;
    double Length() const { return Length(0, 1); }
    double Length(double t1, double t2) const;
    virtual bool device_desc(int& type, double* params4) const
    {
        (void)type;
        (void)params4;
        return false;
    }
    virtual ~Curve() {}
};

typedef std::shared_ptr<Curve> pCurve;

} // namespace bem2d

#endif
