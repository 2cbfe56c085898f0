// Drop-in replacement of lib/bem2d_point.h:11-33 / bem2d_point.cpp:9-61.
#ifndef BEM2D_DROPIN_POINT_H
#define BEM2D_DROPIN_POINT_H

#include <cmath>
#include <ostream>
#include <vector>

#include "bem2d_defs.h"

namespace bem2d {

struct Point {
    double x;
    double y;
    Point(double xx, double yy) : x(xx), y(yy) {}
    Point() : x(0), y(0) {}
};

inline std::ostream& operator<<(std::ostream& s, const Point& z)
{
    return s << "(" << z.x << "," << z.y << ")";
}
inline Point operator+(const Point& p1, const Point& p2) { return Point(p1.x + p2.x, p1.y + p2.y); }
inline Point operator-(const Point& p1, const Point& p2) { return Point(p1.x - p2.x, p1.y - p2.y); }
inline Point operator*(const Point& p, double t) { return Point(t * p.x, t * p.y); }
inline Point operator*(double t, const Point& p) { return Point(t * p.x, t * p.y); }
inline double length(const Point& p) { return std::sqrt(p.x * p.x + p.y * p.y); }
inline Point normalize(const Point& p)
{
    double l = length(p);
    return Point(p.x / l, p.y / l);
}

typedef std::vector<Point> PointVector;
typedef std::shared_ptr<PointVector> pPointVector;

} // namespace bem2d

#endif
