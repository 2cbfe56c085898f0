"""Host description of P0 geometries in the C-ABI's element format, and mesh builders.

A Geometry holds, per element (= matrix row/column), the element type, its 8 parameters and the
index / next / prev adjacency (-1 = none) that selects the singular quadrature rules.  The builders
mirror the reference's element constructions:

  curve_mesh         AnalyticCurveElement(i/n, (i+1)/n, curve) with ring adjacency, as built by hand
                     in the reference's drivers (normcomps/cracknormcond.cpp:67-75) -- uniform in
                     the curve parameter, not by arc length
  analytic_curve     AnalyticCurve(n, curve, closed): n pieces of equal arc length
                     (lib/bem2d_shape.cpp:252-276)
  analytic_curve_ppw AnalyticCurve(ppw, k, curve, closed, L, sigma): ppw pieces per wavelength,
                     the first and last graded (lib/bem2d_shape.cpp:278-331)
  disk_chords        DiskShapePiecewiseConst(n, r)        lib/bem2d_shape.cpp:11-34
  polygon            Polygon(points, n)                   lib/bem2d_shape.cpp:114-138
  line               Line(p0, p1, n) with open ends        lib/bem2d_shape.cpp:225-236
  polygon_ppw        Polygon(points, ppw, k, L, sigma)     lib/bem2d_shape.cpp:36-111
  line_ppw           Line(p0, p1, ppw, k, L, sigma)        lib/bem2d_shape.cpp:147-222
"""
import math

import numpy as np

from . import capi
from .capi import (ELEM_CIRCLE, ELEM_ELLIPSE, ELEM_ELLIPSE_ARC, ELEM_FLAT, ELEM_INV_ELLIPSE,
                   ELEM_KITE, ELEM_TREFOIL)

CURVES = {"circle": ELEM_CIRCLE, "trefoil": ELEM_TREFOIL, "kite": ELEM_KITE,
          "ellipse": ELEM_ELLIPSE, "ellipse_arc": ELEM_ELLIPSE_ARC,
          "inv_ellipse": ELEM_INV_ELLIPSE}


class Geometry:
    def __init__(self, type_, param, index, next_, prev):
        self.type = np.ascontiguousarray(type_, dtype=np.int32)
        self.param = np.ascontiguousarray(param, dtype=np.float64).reshape(-1, 8)
        self.index = np.ascontiguousarray(index, dtype=np.int64)
        self.next = np.ascontiguousarray(next_, dtype=np.int64)
        self.prev = np.ascontiguousarray(prev, dtype=np.int64)
        self.n = int(self.type.size)
        assert self.param.shape[0] == self.n == self.index.size == self.next.size == \
            self.prev.size

    def flat_endpoints(self):
        """For flat elements: (p1, p2) arrays."""
        p1 = self.param[:, 0:2]
        return p1, p1 + self.param[:, 2:4]


def _ring(n, closed=True):
    idx = np.arange(n, dtype=np.int64)
    nxt = (idx + 1) % n
    prv = (idx - 1) % n
    if not closed:
        nxt[-1] = -1
        prv[0] = -1
    return idx, nxt, prv


def curve_mesh(curve, n, params=(0.0, 0.0, 0.0, 0.0), closed=True):
    """n AnalyticCurveElements at uniform curve parameter t_i = i/n."""
    ty = CURVES[curve] if isinstance(curve, str) else int(curve)
    param = np.zeros((n, 8))
    i = np.arange(n, dtype=np.float64)
    param[:, 0] = i / n
    param[:, 1] = (i + 1) / n
    param[:, 2:2 + len(params)] = params
    return Geometry(np.full(n, ty), param, *_ring(n, closed))


def _curve(curve, params):
    ty = CURVES[curve] if isinstance(curve, str) else int(curve)
    args = np.zeros(4)
    args[:len(params)] = params
    return ty, args


def _arclength_mesh(ty, args, partition, closed):
    """Elements between the curve parameters of the arc lengths partition[0 .. n] (partition[0] =
    0); the last element ends at t = 1 (AnalyticCurve::ParameterizeArc sets it)."""
    n = len(partition) - 1
    t = np.zeros(n + 1)
    t[1:] = capi.arclength_params(ty, args, np.asarray(partition[1:], dtype=np.float64))
    t[n] = 1.0
    param = np.zeros((n, 8))
    param[:, 0] = t[:-1]
    param[:, 1] = t[1:]
    param[:, 2:6] = args
    return Geometry(np.full(n, ty), param, *_ring(n, closed))


def analytic_curve(curve, n, params=(0.0, 0.0, 0.0, 0.0), closed=True):
    """AnalyticCurve(n, curve, closed): n elements of equal arc length."""
    ty, args = _curve(curve, params)
    L = capi.curve_length(ty, args)
    return _arclength_mesh(ty, args, [i * L / n for i in range(n + 1)], closed)


def analytic_curve_ppw(curve, ppw, k, params=(0.0, 0.0, 0.0, 0.0), closed=True, L=0, sigma=0.15):
    """AnalyticCurve(ppw, k, curve, closed, L, sigma): c = max(ceil(ppw k len / 2 pi), 10) pieces
    of equal arc length, the first and the last split into L + 1 pieces graded by sigma towards
    the curve's ends."""
    ty, args = _curve(curve, params)
    length = capi.curve_length(ty, args)
    c = max(int(math.ceil(ppw * float(k) * length / 2.0 / math.pi)), 10)
    part = [0.0]
    if L == 0:
        part.append(length / c)
    else:
        ts = 1.0
        for _ in range(L):
            ts = ts * sigma
        part.append(ts * length / c)
        for _ in range(L):
            ts = ts / sigma
            part.append(ts * length / c)
    part += [length / c * j for j in range(2, c)]
    if L == 0:
        part.append(length)
    else:
        td = length / c * (c - 1)
        ts = 1.0
        for _ in range(L):
            ts = ts * sigma
            part.append(td + length / c * (1 - ts))
        part.append(length)
    return _arclength_mesh(ty, args, part, closed)


def disk_chords(n, radius=1.0):
    """DiskShapePiecewiseConst: n flat chords of a circle (lib/bem2d_shape.cpp:11-34)."""
    angle = 2 * math.pi / n
    pts = [(radius * math.cos(i * angle), radius * math.sin(i * angle)) for i in range(n)]
    param = np.zeros((n, 8))
    for i in range(n):
        p1 = pts[i]
        p2 = pts[(i + 1) % n]
        param[i, :4] = (p1[0], p1[1], p2[0] - p1[0], p2[1] - p1[1])
    return Geometry(np.full(n, ELEM_FLAT), param, *_ring(n))


def polygon(points, per_side):
    """Polygon(points, n): n flat elements per side, closed ring (lib/bem2d_shape.cpp:114-138)."""
    m = len(points)
    pts = [tuple(map(float, p)) for p in points] + [tuple(map(float, points[0]))]
    param = np.zeros((m * per_side, 8))
    for s in range(m):
        dx = pts[s + 1][0] - pts[s][0]
        dy = pts[s + 1][1] - pts[s][1]
        for j in range(per_side):
            # p1 = p + (j/n) d, p2 = p + ((j+1)/n) d, stored as p1 and p2 - p1
            p1 = (pts[s][0] + (1.0 * j) / per_side * dx, pts[s][1] + (1.0 * j) / per_side * dy)
            p2 = (pts[s][0] + (1.0 * (j + 1)) / per_side * dx,
                  pts[s][1] + (1.0 * (j + 1)) / per_side * dy)
            param[s * per_side + j, :4] = (p1[0], p1[1], p2[0] - p1[0], p2[1] - p1[1])
    return Geometry(np.full(m * per_side, ELEM_FLAT), param, *_ring(m * per_side))


def _graded_side(params, p, q, n, L, sigma):
    """Pieces of the side p -> q: n of |q - p| / n, the first and the last split into L + 1 pieces
    of ratio sigma (as the ppw-graded Polygon and Line place them)."""
    p = np.asarray(p, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    d = q - p

    def add(a, b):
        params.append([a[0], a[1], b[0] - a[0], b[1] - a[1], 0, 0, 0, 0])

    if L == 0:
        add(p, p + (1.0 / n) * d)
    else:
        ts = 1.0
        for _ in range(L):
            ts = ts * sigma
        add(p, p + (ts / n) * d)
        for _ in range(L):
            a = p + ts / n * d
            ts = ts / sigma
            add(a, p + ts / n * d)
    for j in range(1, n - 1):
        add(p + (1.0 * j) / n * d, p + (1.0 * (j + 1)) / n * d)
    if L == 0:
        add(p + (1.0 * (n - 1)) / n * d, q)
    else:
        ts = 1.0
        for _ in range(L):
            a = p + (1.0 * (n - 1)) / n * d + (1 - ts) / n * d
            ts *= sigma
            add(a, p + (1.0 * (n - 1)) / n * d + (1 - ts) / n * d)
        add(p + (1.0 * (n - 1)) / n * d + (1 - ts) / n * d, q)


def polygon_ppw(points, ppw, k, L=0, sigma=0.15):
    """Polygon(points, ppw, k, L, sigma): max(150, ceil(ppw Re k |side| / 2 pi)) pieces per side,
    graded into the corners; closed ring."""
    pts = [np.asarray(p, dtype=np.float64) for p in points]
    params = []
    for i in range(len(pts)):
        a, b = pts[i], pts[(i + 1) % len(pts)]
        n = max(150, int(math.ceil(ppw * float(complex(k).real) * math.hypot(*(b - a)) / 2.0 /
                                   math.pi)))
        _graded_side(params, a, b, n, L, sigma)
    m = len(params)
    return Geometry(np.full(m, ELEM_FLAT), np.array(params), *_ring(m))


def line_ppw(p0, p1, ppw, k, L=0, sigma=0.15):
    """Line(p0, p1, ppw, k, L, sigma): max(10, ceil(ppw Re k |p1 - p0| / 2 pi)) pieces graded
    towards both ends; the reference links them into a closed ring."""
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    n = max(10, int(math.ceil(ppw * float(complex(k).real) * math.hypot(*(p1 - p0)) / 2.0 /
                              math.pi)))
    params = []
    _graded_side(params, p0, p1, n, L, sigma)
    m = len(params)
    return Geometry(np.full(m, ELEM_FLAT), np.array(params), *_ring(m))


def line(p0, p1, n):
    """Line(p0, p1, n): n flat elements, open ends (-1 sentinels) (lib/bem2d_shape.cpp:225-236)."""
    d = (p1[0] - p0[0], p1[1] - p0[1])
    param = np.zeros((n, 8))
    for j in range(n):
        a = (p0[0] + (1.0 * j) / n * d[0], p0[1] + (1.0 * j) / n * d[1])
        b = (p0[0] + (1.0 * (j + 1)) / n * d[0], p0[1] + (1.0 * (j + 1)) / n * d[1])
        param[j, :4] = (a[0], a[1], b[0] - a[0], b[1] - a[1])
    return Geometry(np.full(n, ELEM_FLAT), param, *_ring(n, closed=False))


# ---- the workloads the benchmarks and parity tests use ---------------------------------------

def unit_circle(n):
    return curve_mesh("circle", n)


def kite(n):
    """Kite x = cos t + 0.65 cos 2t - 0.65, y = 1.5 sin t (lib/bem2d_curvetypes.h:39-48)."""
    return curve_mesh("kite", n)


def open_arc(n, aperture=0.3, radius=1.0):
    """Open circular arc of the given aperture centred on +x: EllipseArc(r, r, a/2, 2 pi - a/2)."""
    return curve_mesh("ellipse_arc", n, (radius, radius, aperture / 2, 2 * math.pi - aperture / 2),
                      closed=False)
