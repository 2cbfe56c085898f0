"""Pins the oracle's Galerkin assembly (oracle_assemble):
- every entry of tiny meshes against mpmath evaluations with the same rule nodes;
- the kernel conventions against analytic eigenvalues of the operators on the unit circle
  (a wrong sign, normal or factor 1/2 changes the row sums by O(1));
- the P0 mass matrix (EvalIdent) against the exact element lengths."""
import numpy as np
import pytest
import scipy.special as sp

from bem2d_amd import geometry as G
from conftest import golden

ENTRY_CASES = ["circle4", "kite5", "triangle6", "arc4open", "circle3_q16"]


def _fixture_geometry(f, name):
    return G.Geometry(f[name + "_type"], f[name + "_param"], f[name + "_index"],
                      f[name + "_next"], f[name + "_prev"])


@pytest.mark.parametrize("name", ENTRY_CASES)
def test_entries_match_mpmath(oracle, name):
    f = golden("entries_mpmath.npz")
    g = _fixture_geometry(f, name)
    N, L, sigma = f[name + "_quad"]
    k, eta = f[name + "_keta"]
    kinds = [int(key.rsplit("kind", 1)[1]) for key in f.files if key.startswith(name + "_kind")]
    assert kinds
    for kind in kinds:
        ref = f["%s_kind%d" % (name, kind)]
        if kind == 0:
            A = oracle.assemble(g, int(N), int(L), sigma, 0, k, eta, 1.0, 0.0)
        else:
            A = oracle.assemble(g, int(N), int(L), sigma, kind, k, eta, 0.0, 1.0)
        err = np.abs(A - ref).max() / np.abs(ref).max()
        assert err < 2e-14, (name, kind, err)


@pytest.mark.parametrize("n,k", [(64, 5.0), (96, 11.0)])
def test_circle_operator_eigenvalues(oracle, n, k):
    g = G.unit_circle(n)
    h = 2 * np.pi / n
    lam = {
        oracle.K_SL: 1j * np.pi / 2 * sp.jv(0, k) * sp.hankel1(0, k),
        oracle.K_DL: -1j * np.pi * k / 2 * sp.jv(0, k) * sp.hankel1(1, k) + 0.5,
        oracle.K_CDL: -1j * np.pi * k / 2 * sp.jv(1, k) * sp.hankel1(0, k) - 0.5,
    }
    tol = {oracle.K_SL: 3e-4, oracle.K_DL: 1e-5, oracle.K_CDL: 1e-5}
    for kind, val in lam.items():
        A = oracle.assemble(g, 16, 5, 0.15, kind, k, k, 0.0, 1.0)
        rows = A.sum(axis=1) / h
        assert np.abs(rows - val).max() / abs(val) < tol[kind], kind
    # combined = CDL - i eta SL, linear in its parts
    eta = 2.5
    Ac = oracle.assemble(g, 16, 3, 0.15, oracle.K_COMBINED, k, eta, 0.0, 1.0)
    As = oracle.assemble(g, 16, 3, 0.15, oracle.K_SL, k, eta, 0.0, 1.0)
    Ad = oracle.assemble(g, 16, 3, 0.15, oracle.K_CDL, k, eta, 0.0, 1.0)
    assert np.abs(Ac - (Ad - 1j * eta * As)).max() < 1e-14 * np.abs(Ac).max()


def test_mass_matrix(oracle):
    n = 40
    M = oracle.assemble(G.unit_circle(n), 16, 3, 0.15, oracle.K_NONE, 1.0, 1.0, 1.0, 0.0)
    np.testing.assert_allclose(np.diag(M).real, 2 * np.pi / n, rtol=1e-14)
    assert np.abs(M - np.diag(np.diag(M))).max() == 0
    tri = G.polygon([(0, 0), (2, 0), (0, 1)], 3)
    M = oracle.assemble(tri, 5, 3, 0.15, oracle.K_NONE, 1.0, 1.0, 1.0, 0.0)
    p1, p2 = tri.flat_endpoints()
    np.testing.assert_allclose(np.diag(M).real, np.hypot(*(p2 - p1).T), rtol=1e-15)


def test_rule_selection_open_curve(oracle):
    """Open arcs: the end elements have one neighbour, so their missing neighbour entry uses the
    regular tensor rule (no special rule is selected through the -1 sentinel)."""
    g = G.open_arc(6)
    A = oracle.assemble(g, 8, 3, 0.15, oracle.K_SL, 4.0, 4.0, 0.0, 1.0)
    assert np.all(np.isfinite(A))
    np.testing.assert_allclose(A, A.T, rtol=1e-12, atol=1e-15)  # SL is symmetric


def test_rhs_flat_closed_form(oracle):
    """On a flat element the 1-D plane-wave integral is analytic:
    int_0^1 exp(ik d.(p1 + t D)) |D| dt = |D| e^{ia} (e^{ib} - 1)/(ib)."""
    tri = G.polygon([(0, 0), (1.3, 0.2), (0.1, 0.8)], 4)
    k = 7.0
    d = np.array([np.cos(0.7), np.sin(0.7)])
    b = oracle.rhs(tri, 16, oracle.F_PLANE, k, 0.0, d, 1.0)[:, 0]
    p1, p2 = tri.flat_endpoints()
    D = p2 - p1
    a = k * p1 @ d
    bb = k * D @ d
    exact = np.hypot(*D.T) * np.exp(1j * a) * (np.exp(1j * bb) - 1) / (1j * bb)
    np.testing.assert_allclose(b, exact, rtol=1e-13)
