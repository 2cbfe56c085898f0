// Drop-in replacement of lib/bem2d_basistypes.h:28-99 and lib/bem2d_basistypes.cpp:10-46.
#ifndef BEM2D_DROPIN_BASISTYPES_H
#define BEM2D_DROPIN_BASISTYPES_H

#include "bem2d_basis.h"
#include "bem2d_geometry.h"

namespace bem2d {

class PolBasis : public Basis {
public:
    static void AddBasis(int deg, pGeometry pgeom)
    {
        for (int i = 0; i <= deg; i++) pgeom->AddBasis(pBasis(new PolBasis(i)));
    }
    explicit PolBasis(int degree) : degree_(degree) {}
    complex operator()(double t) const override
    {
        double result = 1;
        for (int i = 0; i < degree_; i++) result = result * t;
        return result;
    }
    int pol_degree() const override { return degree_; }

private:
    int degree_;
};
typedef std::shared_ptr<PolBasis> pPolBasis;

class LinBasis : public Basis {
public:
    static void AddBasis(pGeometry pgeom)
    {
        pgeom->AddBasis(pBasis(new LinBasis(1)));
        pgeom->AddBasis(pBasis(new LinBasis(-1)));
    }
    explicit LinBasis(int sign) : sign_(sign) {}
    complex operator()(double t) const override { return sign_ == -1 ? 1 - t : t; }

private:
    int sign_;
};
typedef std::shared_ptr<LinBasis> pLinBasis;

class WaveBasis : public Basis {
public:
    static void AddBasis(freqtype k, pGeometry pgeom)
    {
        pgeom->AddBasis(pBasis(new WaveBasis(1.0, k)));
        pgeom->AddBasis(pBasis(new WaveBasis(-1.0, k)));
    }
    WaveBasis(double direction, freqtype k) : direction_(direction), k_(k) {}
    complex operator()(double t) const override
    {
        return std::exp(complex(0, 1) * complex(k_.re, k_.im) * direction_ * t);
    }

private:
    double direction_;
    freqtype k_;
};
typedef std::shared_ptr<WaveBasis> pWaveBasis;

} // namespace bem2d

#endif
