"""Pins the oracle's Gauss rule and singular rules (oracle/bem2d_oracle.c: oracle_gauss,
oracle_rule) against independent facts: mpmath Gauss-Legendre nodes, exactness on monomials, the
covered area of each graded rule and closed-form integrals over that covered region."""
import math

import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 16, 30])
def test_gauss_matches_mpmath(oracle, n):
    g = golden("gauss_mpmath.npz")
    x, w = oracle.gauss(n)
    assert np.all(np.diff(x) > 0) or n == 1  # ascending, as dstev returns them
    np.testing.assert_allclose(x, g["x%d" % n], rtol=0, atol=4e-16)
    np.testing.assert_allclose(w, g["w%d" % n], rtol=0, atol=4e-16)


@pytest.mark.parametrize("n", [5, 16])
def test_gauss_exact_on_monomials(oracle, n):
    x, w = oracle.gauss(n)
    for m in range(2 * n):
        assert abs(np.dot(w, x ** m) - 1.0 / (m + 1)) < 1e-15


def test_gauss_rejects_zero_points(oracle):
    with pytest.raises(ValueError):
        oracle.gauss(0)


def _covered_integral(which, L, s, a, b):
    """Exact integral of x^a y^b over the region each rule covers."""
    def box(x0, x1, y0, y1):
        return (x1 ** (a + 1) - x0 ** (a + 1)) / (a + 1) * (y1 ** (b + 1) - y0 ** (b + 1)) / (b + 1)
    sl = s ** L
    if which == 1:  # [s^L, 1]^2
        return box(sl, 1, sl, 1)
    if which == 2:  # unit square minus the corner box [0, s^L] x [1 - s^L, 1]
        return box(0, 1, 0, 1) - box(0, sl, 1 - sl, 1)
    if which == 3:  # unit square minus [1 - s^L, 1] x [0, s^L]
        return box(0, 1, 0, 1) - box(1 - sl, 1, 0, sl)
    return box(0, 1, 0, 1)


@pytest.mark.parametrize("which,N,L,sigma", [(0, 16, 3, 0.15), (1, 16, 3, 0.15), (2, 16, 3, 0.15),
                                             (3, 16, 3, 0.15), (1, 5, 3, 0.15), (2, 7, 2, 0.3),
                                             (3, 4, 5, 0.2)])
def test_rules_count_weight_and_polynomials(oracle, which, N, L, sigma):
    x, y, w = oracle.rule(which, N, L, sigma)
    expect = {0: N * N, 1: N * N * L * L, 2: 3 * N * N * L, 3: 3 * N * N * L}[which]
    assert x.size == expect
    assert np.all((x > 0) & (x < 1) & (y > 0) & (y < 1) & (w > 0))
    for a, b in [(0, 0), (1, 0), (0, 1), (3, 2), (5, 7)]:
        exact = _covered_integral(which, L, sigma, a, b)
        assert abs(np.dot(w, x ** a * y ** b) - exact) < 2e-15 * max(1.0, abs(exact))


def test_graded_rules_integrate_log_singularities(oracle):
    """AdaptedGauss1 resolves log|x - y| after the Duffy split; AdaptedGauss3 resolves
    log((1 - x) + y) at the corner (1, 0) -- both to near machine precision over the covered
    region, which a plain tensor rule cannot do."""
    N, L, s = 16, 3, 0.15
    sl = s ** L
    x, y, w = oracle.rule(3, N, L, s)
    # int over unit square of log((1-x)+y) minus the excluded corner [1-sl,1]x[0,sl]
    def F(u, v):  # antiderivative of log(u+v) in u then v for the square [0,U]x[0,V]
        def G(z):
            return 0.0 if z == 0 else z * z * math.log(z) / 2 - 3 * z * z / 4
        return G(u + v) - G(u) - G(v) + G(0)
    exact = F(1.0, 1.0) - F(sl, sl)
    assert abs(np.dot(w, np.log((1 - x) + y)) - exact) < 1e-13
    g0 = oracle.rule(0, N, L, s)
    plain = np.dot(g0[2], np.log((1 - g0[0]) + g0[1]))
    assert abs(plain - F(1.0, 1.0)) > 1e-6  # the tensor rule alone is visibly worse
