"""Pins the oracle's H0/H1 (long-double series / Miller-Neumann / asymptotic) against mpmath at
50 digits, the Wronskian and the small-x logarithmic behaviour."""
import numpy as np

from conftest import golden


def test_hankel_matches_mpmath(oracle):
    g = golden("hankel_mpmath.npz")
    x, v = g["x"], g["j0y0j1y1"]
    J0, Y0, J1, Y1 = oracle.hankel(x)
    m0 = np.hypot(v[:, 0], v[:, 1])
    m1 = np.hypot(v[:, 2], v[:, 3])
    assert (np.hypot(J0 - v[:, 0], Y0 - v[:, 1]) / m0).max() < 2e-16
    assert (np.hypot(J1 - v[:, 2], Y1 - v[:, 3]) / m1).max() < 2e-16


def test_wronskian(oracle):
    x = np.logspace(-6, 4, 3001)
    J0, Y0, J1, Y1 = oracle.hankel(x)
    w = (J1 * Y0 - J0 * Y1) * (np.pi * x / 2)
    assert np.abs(w - 1).max() < 1e-14


def test_small_argument_log(oracle):
    x = np.array([1e-8, 1e-6, 1e-4])
    J0, Y0, J1, Y1 = oracle.hankel(x)
    gamma = 0.5772156649015329
    np.testing.assert_allclose(Y0, 2 / np.pi * (np.log(x / 2) + gamma), rtol=1e-8)
    np.testing.assert_allclose(Y1, -2 / (np.pi * x), rtol=1e-6)  # next term ~ x^2 ln x
