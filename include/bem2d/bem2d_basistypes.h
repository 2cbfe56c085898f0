// Drop-in replacement of lib/bem2d_basistypes.h:16-99 and lib/bem2d_basistypes.cpp:10-46:
// ConstBasis, PolBasis, LinBasis, WaveBasis, WavePolBasis with the AddBasis helpers.  Any list of
// these that is the same on every element is assembled on the GPU (bem2d_assemble_basis).
#ifndef BEM2D_DROPIN_BASISTYPES_H
#define BEM2D_DROPIN_BASISTYPES_H

#include "bem2d_basis.h"
#include "bem2d_geometry.h"

namespace bem2d {

class ConstBasis : public Basis {
public:
    ConstBasis() {}
    complex operator()(double) const override { return 1.0; }
    bool device_desc(bem2d_basis& out) const override
    {
        out = bem2d_basis{BEM2D_BASIS_CONST, 0, 0.0, 0.0, 0.0};
        return true;
    }
};
typedef std::shared_ptr<ConstBasis> pConstBasis;

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
    bool device_desc(bem2d_basis& out) const override
    {
        out = bem2d_basis{BEM2D_BASIS_POL, degree_, 0.0, 0.0, 0.0};
        return true;
    }

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
    bool device_desc(bem2d_basis& out) const override
    {
        out = bem2d_basis{BEM2D_BASIS_LIN, sign_ == -1 ? -1 : 1, 0.0, 0.0, 0.0};
        return true;
    }

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
    bool device_desc(bem2d_basis& out) const override
    {
        out = bem2d_basis{BEM2D_BASIS_WAVE, 0, direction_, k_.re, k_.im};
        return true;
    }

private:
    double direction_;
    freqtype k_;
};
typedef std::shared_ptr<WaveBasis> pWaveBasis;

class WavePolBasis : public Basis {
public:
    static void AddBasis(int degree, freqtype k, pGeometry pgeom)
    {
        for (int i = 0; i <= degree; i++) {
            pgeom->AddBasis(pBasis(new WavePolBasis(i, 1.0, k)));
            pgeom->AddBasis(pBasis(new WavePolBasis(i, -1.0, k)));
        }
    }
    WavePolBasis(int degree, double direction, freqtype k)
        : direction_(direction), k_(k), degree_(degree)
    {
    }
    complex operator()(double t) const override
    {
        double result = 1;
        for (int i = 0; i < degree_; i++) result = result * t;
        return std::exp(complex(0, 1) * complex(k_.re, k_.im) * direction_ * t) * result;
    }
    bool device_desc(bem2d_basis& out) const override
    {
        out = bem2d_basis{BEM2D_BASIS_WAVEPOL, degree_, direction_, k_.re, k_.im};
        return true;
    }

private:
    double direction_;
    freqtype k_;
    int degree_;
};
typedef std::shared_ptr<WavePolBasis> pWavePolBasis;

} // namespace bem2d

#endif
