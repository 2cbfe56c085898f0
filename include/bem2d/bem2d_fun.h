// Drop-in replacement of lib/bem2d_fun.h:14-93 (incident-field functors).
#ifndef BEM2D_DROPIN_FUN_H
#define BEM2D_DROPIN_FUN_H

#include "bem2d_defs.h"
#include "bem2d_point.h"

namespace bem2d {

class PlaneWave {
public:
    PlaneWave() {}
    PlaneWave(Point direction, freqtype k) : direction_(direction), k_(k) {}
    inline complex operator()(Point p) const
    {
        complex i(0, 1);
        return std::exp(complex(k_.re, k_.im) * i * (direction_.x * p.x + direction_.y * p.y));
    }
    void SetNormal(Point) {}
    inline freqtype k() const { return k_; }
    inline Point direction() const { return direction_; }

private:
    Point direction_;
    freqtype k_{0, 0};
};

class NormalPlaneWave {
public:
    NormalPlaneWave() {}
    NormalPlaneWave(Point direction, freqtype k) : direction_(direction), k_(k) {}
    inline complex operator()(Point p) const
    {
        complex i(0, 1);
        complex f = std::exp(complex(k_.re, k_.im) * i * (direction_.x * p.x + direction_.y * p.y));
        return i * complex(k_.re, k_.im) * f * (n_.x * direction_.x + n_.y * direction_.y);
    }
    inline void SetNormal(Point normal) { n_ = normal; }
    inline freqtype k() const { return k_; }
    inline Point direction() const { return direction_; }

private:
    Point direction_;
    freqtype k_{0, 0};
    Point n_;
};

class CombinedPlaneWave {
public:
    CombinedPlaneWave(Point direction, freqtype k, double eta)
        : direction_(direction), k_(k), eta_(eta) {}
    CombinedPlaneWave(Point direction, freqtype k) : direction_(direction), k_(k), eta_(k.re) {}
    inline complex operator()(Point p) const
    {
        complex i(0, 1);
        complex f = std::exp(complex(k_.re, k_.im) * i * (direction_.x * p.x + direction_.y * p.y));
        return i * complex(k_.re, k_.im) * f * (n_.x * direction_.x + n_.y * direction_.y) -
               i * eta_ * f;
    }
    inline void SetNormal(Point normal) { n_ = normal; }
    inline freqtype k() const { return k_; }
    inline Point direction() const { return direction_; }
    inline double eta() const { return eta_; }

private:
    Point direction_;
    freqtype k_;
    Point n_;
    double eta_;
};

} // namespace bem2d

#endif
