"""Complex wavenumbers (the kernels' k.im != 0 branch, AMOS zbesh in the reference) on the CPU:
the oracle's and the library's complex Hankel routines against mpmath and against scipy's
hankel1, which is AMOS zbesh itself (lib/bem2d_mathroutines.cpp:44-75); the oracle's complex-k
matrices against entries evaluated entirely in mpmath and against the analytic eigenvalues of the
operators on the unit circle."""
import numpy as np
import pytest
import scipy.special as sp

from bem2d_amd import capi
from bem2d_amd import geometry as G
from conftest import golden


def _rel(got, ref):
    return np.abs(got - ref) / np.abs(ref)


def test_oracle_hankel_complex_matches_mpmath(oracle):
    f = golden("hankel_complex_mpmath.npz")
    z, v = f["z"], f["h0h1"]
    h0, h1 = oracle.hankel_complex(z)
    assert _rel(h0, v[:, 0] + 1j * v[:, 1]).max() < 2e-15
    assert _rel(h1, v[:, 2] + 1j * v[:, 3]).max() < 2e-15


def test_oracle_hankel_complex_matches_amos(oracle):
    rng = np.random.default_rng(5)
    r = np.exp(rng.uniform(np.log(1e-6), np.log(2e3), 4000))
    z = r * np.exp(1j * rng.uniform(-0.95 * np.pi / 2, 0.95 * np.pi / 2, r.size))
    z = z[np.abs(z.imag) < 650]
    h0, h1 = oracle.hankel_complex(z)
    assert _rel(h0, sp.hankel1(0, z)).max() < 1e-13
    assert _rel(h1, sp.hankel1(1, z)).max() < 1e-13


def test_oracle_hankel_complex_identities(oracle):
    # the real axis: the same values as the real-argument routine (a different algorithm)
    x = np.geomspace(1e-3, 1e3, 200)
    h0, h1 = oracle.hankel_complex(x + 0j)
    j0, y0, j1, y1 = oracle.hankel(x)
    assert _rel(h0, j0 + 1j * y0).max() < 1e-14
    assert _rel(h1, j1 + 1j * y1).max() < 1e-14
    # H0' = -H1 (central differences, step 1e-5: truncation ~1e-11, rounding ~1e-11)
    z = np.geomspace(0.5, 200, 60) * np.exp(-0.4j)
    h = 1e-5
    d0 = (oracle.hankel_complex(z + h)[0] - oracle.hankel_complex(z - h)[0]) / (2 * h)
    assert _rel(d0, -oracle.hankel_complex(z)[1]).max() < 1e-8


def test_host_hankel_complex_matches_mpmath():
    """The library's FP64 routine (the device code compiled for the host)."""
    f = golden("hankel_complex_mpmath.npz")
    z, v = f["z"], f["h0h1"]
    phis = np.round(np.angle(z), 12)
    worst = 0.0
    for phi in np.unique(phis):
        sel = phis == phi
        out = capi.hankel01_complex_host(float(phi), z[sel])
        for c in (0, 2):
            ref = v[sel, c] + 1j * v[sel, c + 1]
            worst = max(worst, _rel(out[:, c] + 1j * out[:, c + 1], ref).max())
    assert worst < 1e-14, worst


@pytest.mark.parametrize("name", ["circle4_ck", "triangle6_ck", "kite5_ck"])
def test_oracle_complex_entries_match_mpmath(oracle, name):
    f = golden("entries_complex_mpmath.npz")
    geom = G.Geometry(f[name + "_type"], f[name + "_param"], f[name + "_index"],
                      f[name + "_next"], f[name + "_prev"])
    N, L, sigma = f[name + "_quad"]
    kr, ki, eta = f[name + "_kc"]
    for key in f.files:
        if not key.startswith(name + "_kind"):
            continue
        kind = int(key[len(name) + 5:])
        A = oracle.assemble(geom, int(N), int(L), float(sigma), kind, complex(kr, ki), eta, 0.0,
                            1.0)
        ref = f[key]
        err = np.linalg.norm(A - ref) / np.linalg.norm(ref)
        assert err < 1e-13, (name, kind, err)


@pytest.mark.parametrize("k", [5.0 - 1.0j, 4.0 + 1.5j])
def test_oracle_complex_circle_operator_eigenvalues(oracle, k):
    """Row sums of the P0 matrices on the unit circle approach the operators' eigenvalues for the
    constant density: (i pi/2) J0(k) H0(k) for SL, -(i pi k/2) J0(k) H1(k) + 1/2 for DL,
    -(i pi k/2) J1(k) H0(k) - 1/2 for CDL, with complex k."""
    lam = {
        oracle.K_SL: 1j * np.pi / 2 * sp.jv(0, k) * sp.hankel1(0, k),
        oracle.K_DL: -1j * np.pi * k / 2 * sp.jv(0, k) * sp.hankel1(1, k) + 0.5,
        oracle.K_CDL: -1j * np.pi * k / 2 * sp.jv(1, k) * sp.hankel1(0, k) - 0.5,
    }
    tol = {oracle.K_SL: 5e-3, oracle.K_DL: 5e-4, oracle.K_CDL: 5e-4}
    for kind, val in lam.items():
        err = []
        for n in (32, 64):
            g = G.unit_circle(n)
            A = oracle.assemble(g, 8, 3, 0.15, kind, k, 0.0, 0.0, 1.0)
            rows = A.sum(axis=1) / (2 * np.pi / n)
            err.append(np.abs(rows - val).max() / abs(val))
        # first-order convergence of the P0 discretisation towards the eigenvalue
        assert err[1] < tol[kind] and err[1] < 0.6 * err[0], (kind, k, err)
    g = G.unit_circle(32)
    # a real wavenumber given as complex takes the real-k branch, bit for bit
    Ar = oracle.assemble(g, 8, 3, 0.15, oracle.K_COMBINED, 5.0, 5.0, 1.0, 2.0)
    Ac = oracle.assemble(g, 8, 3, 0.15, oracle.K_COMBINED, 5.0 + 0j, 5.0, 1.0, 2.0)
    assert np.array_equal(Ar, Ac)


def test_oracle_complex_rhs_flat_closed_form(oracle):
    """Plane wave exp(i k d.x) with complex k over one flat element: (e^{ia} - 1)/(ia) * length,
    a = k d.(p1 - p0)... times e^{i k d.p0}."""
    p0, p1 = np.array([0.2, -0.1]), np.array([0.9, 0.4])
    param = np.zeros((1, 8))
    param[0, :4] = [p0[0], p0[1], p1[0] - p0[0], p1[1] - p0[1]]
    geom = G.Geometry(np.zeros(1, np.int32), param, np.zeros(1, np.int64),
                      -np.ones(1, np.int64), -np.ones(1, np.int64))
    k = 7.0 - 2.0j
    d = np.array([np.cos(0.7), np.sin(0.7)])
    b = oracle.rhs(geom, 30, oracle.F_PLANE, k, 0.0, d)[0, 0]
    a = k * d.dot(p1 - p0)
    ref = np.exp(1j * k * d.dot(p0)) * (np.exp(1j * a) - 1) / (1j * a) * np.linalg.norm(p1 - p0)
    assert abs(b - ref) / abs(ref) < 1e-14
