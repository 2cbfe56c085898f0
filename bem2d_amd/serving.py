"""Batched and served workloads: seeded scattering problems of uneven size, and their replay
through the serving engine (bem2d_engine in include/bem2d_hip.h).

  batched_problems(64, seed=1)     ellipses (semi-axes U[0.5, 2]) and polygons (3-8 vertices at
                                   sorted random angles on the unit circle), k log-uniform in
                                   [5, 200], 10 elements per wavelength, one plane wave each
  serving_trace(count, rate, seed=2)
                                   ellipse / kite / polygon, k uniform in [10, 400], 10 elements
                                   per wavelength, 1-32 plane waves at random angles, Poisson
                                   arrivals at the given rate

Element counts follow N = ceil(10 k P / (2 pi)) for perimeter P; polygons get ceil(10 k len / 2 pi)
flat elements per side.  Curves use AnalyticCurveElements at uniform parameter spacing (as the
reference's drivers build them), so elements are not of equal arc length.
"""
import math

import numpy as np

from . import geometry as G
from .capi import ELEM_FLAT, PRECISION_MIXED


def _perimeter(fx, fy, m=4096):
    t = np.linspace(0.0, 1.0, m + 1)
    x, y = fx(t), fy(t)
    return float(np.sum(np.hypot(np.diff(x), np.diff(y))))


def ellipse(a, b, n):
    """EllipseArc(a, b, 0, 2 pi) as a closed ring of n elements."""
    return G.curve_mesh("ellipse_arc", n, (a, b, 0.0, 2 * math.pi))


def polygon_sides(points, counts):
    """Closed polygon with counts[s] flat elements on side s (Polygon of lib/bem2d_shape.cpp:114-138
    with a per-side element count)."""
    m = len(points)
    rows = []
    for s in range(m):
        p, q = points[s], points[(s + 1) % m]
        d = (q[0] - p[0], q[1] - p[1])
        c = counts[s]
        for j in range(c):
            a = (p[0] + (1.0 * j) / c * d[0], p[1] + (1.0 * j) / c * d[1])
            b = (p[0] + (1.0 * (j + 1)) / c * d[0], p[1] + (1.0 * (j + 1)) / c * d[1])
            rows.append((a[0], a[1], b[0] - a[0], b[1] - a[1], 0, 0, 0, 0))
    n = len(rows)
    idx = np.arange(n, dtype=np.int64)
    return G.Geometry(np.full(n, ELEM_FLAT), np.array(rows), idx, (idx + 1) % n, (idx - 1) % n)


def _per_wavelength(k, length, per=10):
    return max(1, int(math.ceil(per * k * length / (2 * math.pi))))


def random_ellipse(rng, k):
    a, b = rng.uniform(0.5, 2.0, size=2)
    P = _perimeter(lambda t: a * np.cos(2 * np.pi * t), lambda t: b * np.sin(2 * np.pi * t))
    return ellipse(a, b, _per_wavelength(k, P)), "ellipse(%.3f,%.3f)" % (a, b)


def random_polygon(rng, k):
    m = int(rng.integers(3, 9))
    ang = np.sort(rng.uniform(0, 2 * np.pi, size=m))
    pts = [(math.cos(t), math.sin(t)) for t in ang]
    counts = [_per_wavelength(k, math.dist(pts[s], pts[(s + 1) % m])) for s in range(m)]
    return polygon_sides(pts, counts), "polygon(%d)" % m


def kite_mesh(k):
    P = _perimeter(lambda t: np.cos(2 * np.pi * t) + 0.65 * np.cos(4 * np.pi * t) - 0.65,
                   lambda t: 1.5 * np.sin(2 * np.pi * t))
    return G.kite(_per_wavelength(k, P)), "kite"


def _dirs(rng, nrhs):
    th = rng.uniform(0, 2 * np.pi, size=nrhs)
    return np.stack([np.cos(th), np.sin(th)], axis=1)


def batched_problems(count=64, seed=1, nang=360, precision=PRECISION_MIXED, tol=1e-12):
    rng = np.random.default_rng(seed)
    theta = 2 * np.pi * np.arange(nang) / nang
    out = []
    for _ in range(count):
        k = float(math.exp(rng.uniform(math.log(5.0), math.log(200.0))))
        geom, name = random_ellipse(rng, k) if rng.uniform() < 0.5 else random_polygon(rng, k)
        out.append({"geom": geom, "name": name, "k": k, "eta": k, "dirs": _dirs(rng, 1),
                    "theta": theta, "precision": precision, "tol": tol})
    return out


def serving_trace(count=2000, rate=50.0, seed=2, nang=36, precision=PRECISION_MIXED, tol=1e-12,
                  max_rhs=32, k_range=(10.0, 400.0)):
    """count requests with Poisson arrivals at rate (1/s); returns (requests, arrival times)."""
    rng = np.random.default_rng(seed)
    theta = 2 * np.pi * np.arange(nang) / nang
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    arrivals -= arrivals[0]
    out = []
    for _ in range(count):
        kind = int(rng.integers(0, 3))
        k = float(rng.uniform(*k_range))
        if kind == 0:
            geom, name = random_ellipse(rng, k)
        elif kind == 1:
            geom, name = kite_mesh(k)
        else:
            geom, name = random_polygon(rng, k)
        nrhs = int(rng.integers(1, max_rhs + 1))
        out.append({"geom": geom, "name": name, "k": k, "eta": k, "dirs": _dirs(rng, nrhs),
                    "theta": theta, "precision": precision, "tol": tol})
    return out, arrivals


def latency_summary(infos):
    lat = np.array([i.t_done - i.t_arrival for i in infos])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
