// Drop-in replacement of the element builders of lib/bem2d_shape.h:15-62 / bem2d_shape.cpp:11-368:
// DiskShapePiecewiseConst(n, r), Polygon(points, n), Line(p0, p1, n) and AnalyticCurve (n pieces
// of equal arc length, or ppw points per wavelength graded towards the ends).  AnalyticCurve's
// arc-length parameterisation comes from the library (bem2d_arclength_params: adaptive
// Gauss-Legendre and Newton's method, to rounding level) instead of GSL's RK2 integration of
// dt/ds to 1e-10; it needs a curve the library knows (Curve::device_desc), ParameterException
// otherwise.  The ppw-graded Polygon and Line place ceil(ppw k |side| / 2 pi) pieces on each side
// (at least 150 per polygon side, 10 per line), the end pieces split into L + 1 pieces graded by
// sigma towards the corners (lib/bem2d_shape.cpp:36-111, 147-222).
#ifndef BEM2D_DROPIN_SHAPE_H
#define BEM2D_DROPIN_SHAPE_H

#include <algorithm>
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

// Pieces of one side p -> q graded towards both ends (the side loop of the ppw-graded Polygon and
// Line builders): n pieces of |q - p| / n, the first and the last split into L + 1 pieces of
// ratio sigma.
inline void graded_side(std::vector<pElement>& out, Point p, Point q, int n, int L, double sigma)
{
    const Point direction = q - p;
    int count = (int)out.size();
    if (L == 0) {
        out.push_back(pElement(new ConstElement(p, p + (1.0 / n) * direction, count++)));
    } else {
        double tsigma = 1;
        for (int m = 0; m < L; m++) tsigma = tsigma * sigma;
        out.push_back(pElement(new ConstElement(p, p + (tsigma / n) * direction, count++)));
        for (int j = 0; j < L; j++) {
            const Point p1 = p + tsigma / n * direction;
            tsigma = tsigma / sigma;
            out.push_back(pElement(new ConstElement(p1, p + tsigma / n * direction, count++)));
        }
    }
    for (int j = 1; j < n - 1; j++)
        out.push_back(pElement(new ConstElement(p + (1.0 * j) / n * direction,
                                                p + (1.0 * (j + 1)) / n * direction, count++)));
    if (L == 0) {
        out.push_back(pElement(new ConstElement(p + (1.0 * (n - 1)) / n * direction, q, count++)));
    } else {
        double tsigma = 1;
        for (int j = 0; j < L; j++) {
            const Point p1 = p + (1.0 * (n - 1)) / n * direction + (1 - tsigma) / n * direction;
            tsigma *= sigma;
            const Point p2 = p + (1.0 * (n - 1)) / n * direction + (1 - tsigma) / n * direction;
            out.push_back(pElement(new ConstElement(p1, p2, count++)));
        }
        out.push_back(pElement(new ConstElement(
            p + (1.0 * (n - 1)) / n * direction + (1 - tsigma) / n * direction, q, count++)));
    }
}

class Polygon {
public:
    Polygon(const std::vector<Point>& points, int ppw, freqtype k, int L = 0, double sigma = 0.15)
    {
        std::vector<Point> p(points);
        p.push_back(p[0]);
        for (std::size_t i = 0; i < points.size(); i++) {
            int n = (int)ceil(ppw * (double)k.re * length(p[i + 1] - p[i]) / 2.0 / PI);
            n = std::max(150, n);
            graded_side(elements_, p[i], p[i + 1], n, L, sigma);
        }
        DiskShapePiecewiseConst::link_ring(elements_);
    }
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
    // (the reference closes this graded line into a ring: element 0's prev is the last element)
    Line(Point p0, Point p1, int ppw, freqtype k, int L = 0, double sigma = 0.15)
    {
        int n = (int)ceil(ppw * (double)k.re * length(p1 - p0) / 2.0 / PI);
        n = std::max(10, n);
        graded_side(elements_, p0, p1, n, L, sigma);
        DiskShapePiecewiseConst::link_ring(elements_);
    }
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
