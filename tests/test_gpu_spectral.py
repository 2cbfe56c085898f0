"""Spectral drivers on the GPU (ChangeBasis, SingularValues, L2NormCond, HermitianEigenvalues,
NumRange of lib/bem2d_mathroutines.cpp) through the C-ABI, on matrices assembled by the oracle:
against numpy's textbook routines, and the numerical range against the convex geometry that fixes
it (support lines, and the convex hull of the eigenvalues for a normal matrix)."""
import numpy as np
import pytest
import torch

from bem2d_amd import geometry as G
from conftest import rel_fro

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.asarray(a, dtype=np.complex128)
    t = torch.empty((a.shape[1], a.shape[0]), dtype=torch.complex128, device="cuda")
    t.copy_(torch.from_numpy(np.ascontiguousarray(a.T)))
    return t.t()


def _np(t):
    return t.detach().cpu().numpy()


def _operators(oracle, n=150, k=10.0):
    """The numerical-range drivers' matrices (numrangecomps/disk.cpp:92-95 with P0 elements):
    combined = Id + 2 CDL - 2 i eta SL and the mass matrix Id, on the unit circle."""
    geom = G.unit_circle(n)
    cdl = oracle.assemble(geom, 16, 3, 0.15, oracle.K_CDL, k, k, 0.0, 1.0)
    sl = oracle.assemble(geom, 16, 3, 0.15, oracle.K_SL, k, k, 0.0, 1.0)
    Id = oracle.assemble(geom, 16, 3, 0.15, oracle.K_NONE, k, k, 1.0, 0.0)
    return np.asarray(Id + 2.0 * cdl - 2j * k * sl), np.asarray(Id)


def test_change_basis_against_numpy(ctx, oracle):
    K, Id = _operators(oracle)
    n = K.shape[0]
    rng = np.random.default_rng(4)
    C = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for M in (Id, C @ C.conj().T / n + np.eye(n)):  # the P0 mass matrix, and a dense HPD one
        L = np.linalg.cholesky(M)                      # M = L L^H, so U = L^H
        ref = np.linalg.solve(L, np.linalg.solve(L, K.conj().T).conj().T)  # L^{-1} K L^{-H}
        got = _np(ctx.change_basis(_dev(K), _dev(M)))
        print("change of basis: rel err %.2e" % rel_fro(got, ref))
        assert rel_fro(got, ref) < 1e-13


def test_singular_values_and_norm_cond(ctx, oracle):
    K, Id = _operators(oracle)
    Kd = _dev(K)
    s = _np(ctx.singular_values(Kd))
    ref = np.linalg.svd(K, compute_uv=False)
    assert np.abs(s - ref).max() <= 1e-13 * ref[0]
    assert np.all(np.diff(s) <= 0)
    nrm, cond = ctx.l2_norm_cond(Kd)
    assert abs(nrm - ref[0]) <= 1e-13 * ref[0] and abs(cond - ref[0] / ref[-1]) <= 1e-11 * cond
    print("norm %.6f cond %.6f" % (nrm, cond))


def test_hermitian_eigenvalues(ctx, oracle):
    K, _ = _operators(oracle)
    H = 0.5 * (K + K.conj().T)
    w = _np(ctx.hermitian_eigenvalues(_dev(H)))
    ref = np.linalg.eigvalsh(H)
    assert np.abs(w - ref).max() <= 1e-13 * np.abs(ref).max()


def test_numerical_range_support_lines(ctx, oracle):
    """Every boundary point lies on the support lines of its two angles, with the extreme
    eigenvalues of (e^{it}A + (e^{it}A)^H)/2 from numpy; and the polygon is circumscribed: for
    angles off the grid the support function of the points bounds numpy's from above, by at
    most the polygon's O(step^2) excess."""
    K, Id = _operators(oracle, n=120)
    m = 48
    z = ctx.numerical_range(_dev(K), m)
    assert z.shape == (2 * m + 1,) and z[0] == z[-1]
    th = np.arange(m) * np.pi / m

    def lam(t):
        return np.linalg.eigvalsh(0.5 * (np.exp(1j * t) * K + (np.exp(1j * t) * K).conj().T))

    ext = [lam(t) for t in th]
    scale = max(np.abs(e).max() for e in ext)
    for i in range(m):  # minima at th_i, the first half of the points
        assert abs((np.exp(1j * th[i]) * z[i]).real - ext[i][0]) <= 1e-12 * scale
        assert abs((np.exp(1j * th[i]) * z[m + i]).real - ext[i][-1]) <= 1e-12 * scale
    for t in np.random.default_rng(1).uniform(0, 2 * np.pi, 20):
        top = lam(t)[-1]
        h = (np.exp(1j * t) * z).real.max()
        assert top - 1e-12 * scale <= h <= top + 0.05 * scale


def test_numerical_range_of_a_normal_matrix(ctx):
    """For a normal matrix the numerical range is the convex hull of its eigenvalues: here a
    regular 7-gon of eigenvalues (the others inside it), rotated into a dense unitary frame.  Every
    computed vertex lies outside or on the 7-gon, and within the circumscribed polygon's excess."""
    rng = np.random.default_rng(2)
    n, m = 60, 180
    lam = np.exp(2j * np.pi * np.arange(7) / 7 + 0.3j) * 2.0 + (0.5 - 0.25j)
    d = np.concatenate([lam, lam[rng.integers(0, 7, n - 7)] * rng.uniform(0, 1, n - 7)])
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q * d) @ Q.conj().T
    z = ctx.numerical_range(_dev(A), m)
    # signed distance of each point from the 7-gon: max over edges of the outward offset
    c = 0.5 - 0.25j
    apothem = 2.0 * np.cos(np.pi / 7)
    normals = np.exp(1j * (2 * np.pi * (np.arange(7) + 0.5) / 7 + 0.3))
    off = np.array([max(((zz - c) * np.conj(nv)).real for nv in normals) - apothem for zz in z])
    assert off.min() >= -1e-10          # outside or on the hull
    # two support lines spaced pi/m through the ends of an edge of length L meet at most
    # L tan(pi/(2m)) / 2 beyond it
    edge = 2 * 2.0 * np.sin(np.pi / 7)
    assert off.max() <= edge * np.tan(np.pi / (2 * m)) / 2 + 1e-10
