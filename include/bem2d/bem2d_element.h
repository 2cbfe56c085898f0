// Drop-in replacement of lib/bem2d_element.h:14-116.
// Differences: next/prev start as "none" (-1 as size_t) instead of uninitialised
// (lib/bem2d_element.h:62-65), so open curves need no explicit sentinels; device_desc() (new)
// reports the element in the C-ABI's format.
#ifndef BEM2D_DROPIN_ELEMENT_H
#define BEM2D_DROPIN_ELEMENT_H

#include <cstddef>

#include "../bem2d_hip.h"
#include "bem2d_curve.h"
#include "bem2d_point.h"

namespace bem2d {

class Element {
public:
    Element() = default;
    Element(const Element&) = delete;
    Element& operator=(const Element&) = delete;

    virtual const Point Map(double t) const = 0;
    virtual const Point Deriv(double t) const = 0;
    virtual const double Length() const = 0;

    inline const Point First() { return Map(0); }
    inline const Point Last() { return Map(1); }
    inline const Point Normal(double t) const
    {
        Point p(Deriv(t));
        Point pn(p.y, -p.x);
        return normalize(pn);
    }
    inline void set_index(std::size_t index) { index_ = index; }
    inline std::size_t index() const { return index_; }
    inline void set_next(std::size_t index) { next_ = index; }
    inline std::size_t next() const { return next_; }
    inline void set_prev(std::size_t index) { prev_ = index; }
    inline std::size_t prev() const { return prev_; }

    // Element type and 8 parameters as read by bem2d_mesh_create; false if the GPU cannot
    // evaluate this element by itself.
    virtual bool device_desc(int& type, double* param8) const = 0;
    virtual ~Element() {}

private:
    std::size_t index_ = 0;
    std::size_t next_ = (std::size_t)-1;
    std::size_t prev_ = (std::size_t)-1;
};

typedef std::shared_ptr<Element> pElement;

class ConstElement : public Element {
public:
    ConstElement(const Point& p1, const Point& p2, std::size_t index) : p1_(p1), d_(p2 - p1)
    {
        set_index(index);
    }
    inline const Point Map(double t) const override { return p1_ + d_ * t; }
    inline const Point Deriv(double) const override { return d_; }
    inline const double Length() const override { return length(d_); }
    bool device_desc(int& type, double* p) const override
    {
        type = BEM2D_ELEM_FLAT;
        p[0] = p1_.x;
        p[1] = p1_.y;
        p[2] = d_.x;
        p[3] = d_.y;
        p[4] = p[5] = p[6] = p[7] = 0;
        return true;
    }

private:
    const Point p1_;
    const Point d_;
};

class AnalyticCurveElement : public Element {
public:
    AnalyticCurveElement(double t1, double t2, pCurve curve, std::size_t index)
        : curve_(curve), tstart_(t1), tend_(t2)
    {
        set_index(index);
    }
    inline const Point Map(double t) const override
    {
        return curve_->Map(tstart_ + t * (tend_ - tstart_));
    }
    inline const Point Deriv(double t) const override
    {
        return (tend_ - tstart_) * curve_->Deriv(tstart_ + t * (tend_ - tstart_));
    }
    inline const double Length() const override { return curve_->Length(tstart_, tend_); }
    bool device_desc(int& type, double* p) const override
    {
        if (!curve_->device_desc(type, p + 2)) return false;
        p[0] = tstart_;
        p[1] = tend_;
        p[6] = p[7] = 0;
        return true;
    }

private:
    pCurve curve_;
    double tstart_;
    double tend_;
};

} // namespace bem2d

#endif
