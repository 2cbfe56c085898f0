// Drop-in replacement of the element builders of lib/bem2d_shape.h:15-62 / bem2d_shape.cpp:11-368:
// DiskShapePiecewiseConst(n, r), Polygon(points, n), Line(p0, p1, n) and AnalyticCurve (n pieces
// of equal arc length, or ppw points per wavelength graded towards the ends).  AnalyticCurve's
// arc-length parameterisation comes from the library (bem2d_arclength_params: adaptive
// Gauss-Legendre and Newton's method, to rounding level) instead of GSL's RK2 integration of
// dt/ds to 1e-10; it needs a curve the library knows (Curve::device_desc), ParameterException
// otherwise.  (The ppw-graded Polygon/Line builders are not part of this build.)
#ifndef BEM2D_DROPIN_SHAPE_H
#define BEM2D_DROPIN_SHAPE_H

#include <cmath>
#include <vector>

#include "bem2d_curve.h"
#include "bem2d_element.h"
#include "bem2d_geometry.h"

namespace bem2d {

class DiskShapePiecewiseConst {
public:
    DiskShapePiecewiseConst(int n, double radius) : elements_(n)
    {
        double angle = 2 * PI / n;
        std::vector<Point> pts(n);
        for (int i = 0; i < n; i++) pts[i] = Point(radius * cos(i * angle), radius * sin(i * angle));
        for (int i = 0; i < n; i++)
            elements_[i] = pElement(new ConstElement(pts[i], pts[(i + 1) % n], i));
        link_ring(elements_);
    }
    pGeometry GetGeometry() { return pGeometry(new Geometry(elements_)); }

    static void link_ring(std::vector<pElement>& e)
    {
        const int n = (int)e.size();
        for (int i = 0; i < n; i++) {
            e[i]->set_next((i + 1) % n);
            e[i]->set_prev((i - 1 + n) % n);
        }
    }

private:
    std::vector<pElement> elements_;
};

class Polygon {
public:
    Polygon(const std::vector<Point>& points, int n)
    {
        std::vector<Point> p(points);
        p.push_back(p[0]);
        for (std::size_t i = 0; i < points.size(); i++) {
            Point direction = p[i + 1] - p[i];
            for (int j = 0; j < n; j++) {
                Point p1 = p[i] + (1.0 * j) / n * direction;
                Point p2 = p[i] + (1.0 * (j + 1)) / n * direction;
                elements_.push_back(pElement(new ConstElement(p1, p2, n * i + j)));
            }
        }
        DiskShapePiecewiseConst::link_ring(elements_);
    }
    pGeometry GetGeometry() { return pGeometry(new Geometry(elements_)); }

private:
    std::vector<pElement> elements_;
};

class AnalyticCurve {
public:
    AnalyticCurve(int n, pCurve curve, int closed = 1);
    AnalyticCurve(int ppw, freqtype k, pCurve curve, int closed = 1, int L = 0,
                  double sigma = 0.15);
    pGeometry GetGeometry();
    void ParameterizeArc(dvector& partition);

private:
    void build(int closed);
    std::vector<pElement> elements_;
    pCurve curve_;
    dvector arclengthparam;
};

class Line {
public:
    Line(Point p0, Point p1, int n)
    {
        Point direction = p1 - p0;
        for (int j = 0; j < n; j++) {
            Point pp1 = p0 + (1.0 * j) / n * direction;
            Point pp2 = p0 + (1.0 * (j + 1)) / n * direction;
            elements_.push_back(pElement(new ConstElement(pp1, pp2, j)));
        }
        for (int i = 0; i < n; i++) {
            if (i + 1 < n) elements_[i]->set_next(i + 1);
            if (i > 0) elements_[i]->set_prev(i - 1);
        }
    }
    pGeometry GetGeometry() { return pGeometry(new Geometry(elements_)); }

private:
    std::vector<pElement> elements_;
};

} // namespace bem2d

#endif
