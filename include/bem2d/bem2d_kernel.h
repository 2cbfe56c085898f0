// Drop-in replacement of lib/bem2d_kernel.h:12-98 / bem2d_kernel.cpp:7-112.
// The functors keep the reference's interface (SetNormal + operator()), evaluated on the host with
// the library's FP64 Hankel routine; DiscreteKernel (bem2d_quadrature.h) dispatches them to the GPU.
// Complex wavenumbers (k.im != 0) take the reference's complex branches with the library's complex
// Hankel routine (BesselH0/H1(complex), the stand-in for AMOS zbesh), on the host here and in the
// GPU operators; they need k.re > 0 (ParameterException otherwise).
#ifndef BEM2D_DROPIN_KERNEL_H
#define BEM2D_DROPIN_KERNEL_H

#include "bem2d_defs.h"
#include "bem2d_exceptions.h"
#include "bem2d_point.h"

namespace bem2d {

complex BesselH0(double x);
complex BesselH1(double x);
complex BesselH0(complex x); // Re x > 0; ParameterException otherwise
complex BesselH1(complex x);

class SingleLayer {
public:
    SingleLayer(freqtype k) : k_(k) {}
    complex operator()(Point x, Point y) const
    {
        double d = length(x - y);
        if (k_.im != 0) return complex(0, 1) / 4.0 * BesselH0(complex(k_.re, k_.im) * d);
        return complex(0, 1) / 4.0 * BesselH0(k_.re * d);
    }
    inline void SetNormal(Point normal1, Point normal2)
    {
        n1_ = normal1;
        n2_ = normal2;
    }
    inline freqtype k() const { return k_; }

private:
    freqtype k_;
    Point n1_, n2_;
};

class DoubleLayer {
public:
    DoubleLayer(freqtype k) : k_(k) {}
    complex operator()(Point x, Point y) const
    {
        Point w = y - x;
        double d = length(w);
        if (k_.im != 0) {
            complex kc(k_.re, k_.im);
            return -complex(0, 1) * kc / 4.0 * (BesselH1(kc * d) / d * (n2_.x * w.x + n2_.y * w.y));
        }
        return -complex(0, 1) * k_.re / 4.0 * (BesselH1(k_.re * d) / d * (n2_.x * w.x + n2_.y * w.y));
    }
    inline void SetNormal(Point normal1, Point normal2)
    {
        n1_ = normal1;
        n2_ = normal2;
    }
    inline freqtype k() const { return k_; }

private:
    freqtype k_;
    Point n1_, n2_;
};

class ConjDoubleLayer {
public:
    ConjDoubleLayer(freqtype k) : k_(k) {}
    complex operator()(Point x, Point y) const
    {
        Point w = x - y;
        double d = length(w);
        if (k_.im != 0) {
            complex kc(k_.re, k_.im);
            return -complex(0, 1) * kc / 4.0 * (BesselH1(kc * d) / d * (n1_.x * w.x + n1_.y * w.y));
        }
        return -complex(0, 1) * k_.re / 4.0 * (BesselH1(k_.re * d) / d * (n1_.x * w.x + n1_.y * w.y));
    }
    inline void SetNormal(Point normal1, Point normal2)
    {
        n1_ = normal1;
        n2_ = normal2;
    }
    inline freqtype k() const { return k_; }

private:
    freqtype k_;
    Point n1_, n2_;
};

class CombinedSingleConjDouble {
public:
    CombinedSingleConjDouble(freqtype k, double eta) : k_(k), eta_(eta) {}
    CombinedSingleConjDouble(freqtype k) : k_(k), eta_(k.re) {}
    complex operator()(Point x, Point y) const
    {
        Point w = x - y;
        double d = length(w);
        complex i(0, 1);
        if (k_.im != 0) {
            complex kc(k_.re, k_.im);
            complex slayer = i / 4.0 * BesselH0(kc * d);
            complex cdlayer = -i * kc / 4.0 * (BesselH1(kc * d) / d * (n1_.x * w.x + n1_.y * w.y));
            return cdlayer - i * eta_ * slayer;
        }
        complex slayer = i / 4.0 * BesselH0(k_.re * d);
        complex cdlayer = -i * k_.re / 4.0 * (BesselH1(k_.re * d) / d * (n1_.x * w.x + n1_.y * w.y));
        return cdlayer - i * eta_ * slayer;
    }
    inline void SetNormal(Point normal1, Point normal2)
    {
        n1_ = normal1;
        n2_ = normal2;
    }
    inline freqtype getk() const { return k_; }
    inline double eta() const { return eta_; }

private:
    freqtype k_;
    Point n1_, n2_;
    double eta_;
};

class IdKernel {
public:
    inline void SetNormal(Point, Point) {}
    complex operator()(Point, Point) const { return 1.0; }
};

} // namespace bem2d

#endif
