// Drop-in replacement of lib/bem2d_curvetypes.h:13-99 (the same parameterisations on t in [0, 1],
// with the reference's evaluation order of each product).
#ifndef BEM2D_DROPIN_CURVETYPES_H
#define BEM2D_DROPIN_CURVETYPES_H

#include <cmath>

#include "../bem2d_hip.h"
#include "bem2d_curve.h"

namespace bem2d {

class Circle : public Curve {
public:
    Point Map(double t) const override { return Point(cos(2 * PI * t), sin(2 * PI * t)); }
    Point Deriv(double t) const override
    {
        return 2 * PI * Point(-1.0 * sin(2 * PI * t), cos(2 * PI * t));
    }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_CIRCLE;
        p[0] = p[1] = p[2] = p[3] = 0;
        return true;
    }
};

class Trefoil : public Curve {
public:
    Point Map(double t) const override
    {
        double r = 1 + 0.3 * cos(3 * 2 * PI * t);
        return Point(r * cos(2 * PI * t), r * sin(2 * PI * t));
    }
    Point Deriv(double t) const override
    {
        double r = 1 + 0.3 * cos(3 * 2 * PI * t);
        double rp = -0.3 * 3 * 2 * PI * sin(3 * 2 * PI * t);
        return Point(rp * cos(2 * PI * t) - r * 2 * PI * sin(2 * PI * t),
                     rp * sin(2 * PI * t) + r * 2 * PI * cos(2 * PI * t));
    }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_TREFOIL;
        p[0] = p[1] = p[2] = p[3] = 0;
        return true;
    }
};

class Kite : public Curve {
public:
    Point Map(double t) const override
    {
        return Point(cos(2 * PI * t) + 0.65 * cos(4 * PI * t) - 0.65, 1.5 * sin(2 * PI * t));
    }
    Point Deriv(double t) const override
    {
        return Point(-2 * PI * sin(2 * PI * t) - 0.65 * 4 * PI * sin(4 * PI * t),
                     1.5 * 2 * PI * cos(2 * PI * t));
    }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_KITE;
        p[0] = p[1] = p[2] = p[3] = 0;
        return true;
    }
};

class Ellipse : public Curve {
public:
    explicit Ellipse(double rho) : rho_(rho) {}
    Point Map(double t) const override
    {
        return Point(rho_ * cos(2 * PI * t), 1 / rho_ * sin(2 * PI * t));
    }
    Point Deriv(double t) const override
    {
        return Point(-rho_ * 2 * PI * sin(2 * PI * t), 2 * PI / rho_ * cos(2 * PI * t));
    }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_ELLIPSE;
        p[0] = rho_;
        p[1] = p[2] = p[3] = 0;
        return true;
    }

private:
    double rho_;
};

class EllipseArc : public Curve {
public:
    EllipseArc(double a, double b, double theta1, double theta2)
        : theta1_(theta1), theta2_(theta2), a_(a), b_(b)
    {
    }
    Point Map(double t) const override
    {
        return Point(a_ * cos(theta1_ + t * (theta2_ - theta1_)),
                     b_ * sin(theta1_ + t * (theta2_ - theta1_)));
    }
    Point Deriv(double t) const override
    {
        return Point(-a_ * (theta2_ - theta1_) * sin(theta1_ + t * (theta2_ - theta1_)),
                     b_ * (theta2_ - theta1_) * cos(theta1_ + t * (theta2_ - theta1_)));
    }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_ELLIPSE_ARC;
        p[0] = a_;
        p[1] = b_;
        p[2] = theta1_;
        p[3] = theta2_;
        return true;
    }

private:
    double theta1_, theta2_, a_, b_;
};

class InvEllipse : public Curve {
public:
    explicit InvEllipse(double alpha) : alpha_(alpha) {}
    Point Map(double t) const override
    {
        complex s = 2.0 * PI * t * complex(0, 1);
        complex res = exp(s) / (1.0 + alpha_ * exp(2.0 * s));
        return Point(real(res), imag(res));
    }
    Point Deriv(double t) const override
    {
        complex s = 2.0 * PI * t * complex(0, 1);
        complex i(0, 1);
        complex res = 2.0 * PI * i * exp(s) / (1.0 + alpha_ * exp(2.0 * s)) /
                      (1.0 + alpha_ * exp(2.0 * s)) * (1.0 - alpha_ * exp(2.0 * s));
        return Point(real(res), imag(res));
    }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_INV_ELLIPSE;
        p[0] = alpha_;
        p[1] = p[2] = p[3] = 0;
        return true;
    }

private:
    double alpha_;
};

} // namespace bem2d

#endif
