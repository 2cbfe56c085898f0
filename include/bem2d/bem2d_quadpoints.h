// Drop-in replacement of lib/bem2d_quadpoints.h:12-197: the quadrature rules on [0,1] and [0,1]^2.
// The node lists come from the C-ABI (bem2d_gauss_legendre / bem2d_quad_rule) -- the same tables
// the device assembly uses.
#ifndef BEM2D_DROPIN_QUADPOINTS_H
#define BEM2D_DROPIN_QUADPOINTS_H

#include "../bem2d_hip.h"
#include "bem2d_defs.h"
#include "bem2d_exceptions.h"

namespace bem2d {

struct QuadPoints1D {
    dvector x, w;
    double a, b;
};

struct QuadPoints2D {
    dvector x, y, w;
    double a, b, c, d;
};

class Gauss1D {
public:
    explicit Gauss1D(int N) : N_(N)
    {
        if (N < 1) throw ParameterException();
        x_.resize(N);
        w_.resize(N);
        bem2d_gauss_legendre(N, x_.data(), w_.data());
    }
    QuadPoints1D Rescale(double a, double b) const
    {
        if (a >= b) throw IntervalException();
        QuadPoints1D p{x_, w_, a, b};
        for (std::size_t i = 0; i < p.x.size(); i++) {
            p.x[i] = a + (b - a) * p.x[i];
            p.w[i] = (b - a) * p.w[i];
        }
        return p;
    }
    inline const dvector& x() const { return x_; }
    inline const dvector& w() const { return w_; }
    inline int Size() const { return N_; }

private:
    dvector x_, w_;
    int N_;
};

// Shared storage of the 2-D rules.
class Rule2D {
public:
    inline const dvector& x() const { return x_; }
    inline const dvector& y() const { return y_; }
    inline const dvector& w() const { return w_; }
    inline int Size() const { return N_; }

protected:
    void load(int which, int N, int L, double sigma)
    {
        if (N < 1) throw ParameterException();
        int m = bem2d_quad_rule(which, N, L, sigma, nullptr, nullptr, nullptr);
        if (m < 0) throw ParameterException();
        x_.resize(m);
        y_.resize(m);
        w_.resize(m);
        if (m) bem2d_quad_rule(which, N, L, sigma, x_.data(), y_.data(), w_.data());
        N_ = m;
    }
    dvector x_, y_, w_;
    int N_ = 0;
};

class Gauss2D : public Rule2D {
public:
    explicit Gauss2D(int N) { load(0, N, 1, 0.5); }
    QuadPoints2D Rescale(double a, double b, double c, double d) const
    {
        if (a >= b || c >= d) throw IntervalException();
        QuadPoints2D p{x_, y_, w_, a, b, c, d};
        for (std::size_t i = 0; i < p.x.size(); i++) {
            p.x[i] = a + (b - a) * p.x[i];
            p.y[i] = c + (d - c) * p.y[i];
            p.w[i] = (b - a) * (d - c) * p.w[i];
        }
        return p;
    }
};

class AdaptedGauss1 : public Rule2D {
public:
    AdaptedGauss1(int N, int L, double sigma) { load(1, N, L, sigma); }
};

class AdaptedGauss2 : public Rule2D {
public:
    AdaptedGauss2(int N, int L, double sigma) { load(2, N, L, sigma); }
};

class AdaptedGauss3 : public Rule2D {
public:
    AdaptedGauss3(int N, int L, double sigma) { load(3, N, L, sigma); }
};

} // namespace bem2d

#endif
