"""Several basis functions per element on the GPU (bem2d_assemble_basis, bem2d_rhs_basis,
bem2d_eval_single_layer_basis, bem2d_farfield_basis) against the oracle, and the C++ drop-in driver
shaped like the reference's examples/disk.cpp (AnalyticCurve, LinBasis::AddBasis, SoundSoftScattering,
set_polygons / set_plotInterior, EvalSol, far field)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from bem2d_amd import capi
from bem2d_amd import geometry as G
from conftest import ROOT, analytic_farfield, rel_fro

pytestmark = pytest.mark.gpu

LIN = capi.lin_bases()
POL2 = capi.pol_bases(2)
KINDS = {"sl": 1, "dl": 2, "cdl": 3, "combined": 4, "id": 5}


def _np(t):
    return t.detach().cpu().numpy()


def _tol(oracle, geom, bases, N, L, sigma, kind, k, eta, ms, ks):
    """max(1e-12, 3 x the oracle's rounding noise floor), the floor sampled on 6 columns (the
    correctly rounded trig variant of the oracle is slow)."""
    nn = geom.n * len(bases)
    cols = np.unique(np.linspace(0, nn - 1, 6).astype(np.int64))
    return max(1e-12, 3 * oracle.noise_floor_basis(geom, bases, N, L, sigma, kind, k, eta, ms, ks,
                                                   cols=cols))


CASES = [
    # name, geometry, bases, (N, L, sigma), kind, k, eta, mass_scale, kernel_scale
    ("circle_lin_combined", G.analytic_curve("circle", 60), LIN, (16, 3, 0.15), 4, 8.0, 8.0, 1.0,
     2.0),
    ("kite_pol2_cdl", G.analytic_curve("kite", 50), POL2, (12, 3, 0.15), 3, 6.0, 0.0, 0.0, 1.0),
    ("kite_pol2_dl", G.analytic_curve("kite", 50), POL2, (12, 3, 0.15), 2, 6.0, 0.0, 0.0, 1.0),
    ("polygon_lin_sl", G.polygon([(0, 0), (1.2, 0.1), (0.3, 0.9)], 12), LIN, (8, 2, 0.2), 1, 5.0,
     0.0, 0.0, 1.0),
    ("arc_wave_combined", G.open_arc(40, 0.6), capi.wave_bases(3.0), (16, 3, 0.15), 4, 9.0, 9.0,
     1.0, 2.0),
    ("circle_wavepol1_sl", G.unit_circle(32), capi.wavepol_bases(1, 2.0), (10, 2, 0.15), 1, 4.0,
     0.0, 0.0, 1.0),
    ("trefoil_wavepol2_cdl", G.curve_mesh("trefoil", 30), capi.wavepol_bases(2, 1.5 - 0.2j),
     (8, 2, 0.15), 3, 5.0, 0.0, 0.0, 1.0),
    ("circle_pol4_combined", G.unit_circle(24), capi.pol_bases(4), (8, 2, 0.15), 4, 3.0, 3.0, 1.0,
     2.0),
    ("kite_wavepol3_sl", G.kite(20), capi.wavepol_bases(3, 2.5), (8, 2, 0.15), 1, 4.0, 0.0, 0.0,
     1.0),
    ("kite_lin_mass", G.kite(40), LIN, (16, 3, 0.15), 0, 1.0, 0.0, 1.0, 0.0),
    ("circle_pol2_id", G.unit_circle(20), POL2, (8, 3, 0.15), 5, 1.0, 0.0, 0.0, 1.0),
    ("disk_lin_N30_L7", G.analytic_curve("circle", 40), LIN, (30, 7, 0.15), 4, 6.0, 6.0, 1.0,
     2.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_assemble_basis_matches_oracle(ctx, oracle, case):
    name, geom, bases, (N, L, sigma), kind, k, eta, ms, ks = case
    print("case", name, "n", geom.n, "nb", len(bases), flush=True)
    mesh = ctx.mesh(geom, N, L, sigma)
    A = _np(ctx.assemble_basis(mesh, bases, kind, k, eta, ms, ks))
    ref = oracle.assemble_basis(geom, bases, N, L, sigma, kind, k, eta, ms, ks)
    err = rel_fro(A, ref)
    tol = _tol(oracle, geom, bases, N, L, sigma, kind, k, eta, ms, ks)
    print("%s: rel err %.2e (tol %.2e)" % (name, err, tol), flush=True)
    assert err < tol


@pytest.mark.parametrize("kind", [1, 3, 4])
def test_assemble_basis_complex_k(ctx, oracle, kind):
    geom = G.analytic_curve("kite", 30)
    k = 6.0 - 0.7j
    for bases in (LIN, capi.wave_bases(2.0 - 0.3j)):
        mesh = ctx.mesh(geom, 10, 2, 0.15)
        A = _np(ctx.assemble_basis(mesh, bases, kind, k, 6.0, 1.0, 2.0))
        ref = oracle.assemble_basis(geom, bases, 10, 2, 0.15, kind, k, 6.0, 1.0, 2.0)
        err = rel_fro(A, ref)
        print("complex k kind %d nb %d: %.2e" % (kind, len(bases), err))
        assert err < 1e-12


def test_p0_basis_equals_p0_path(ctx):
    """ConstBasis through the basis kernels equals the P0 kernels (rounding level)."""
    geom = G.kite(300)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    a = _np(ctx.assemble(mesh, 4, 40.0, 40.0, 1.0, 2.0))
    b = _np(ctx.assemble_basis(mesh, capi.const_basis(), 4, 40.0, 40.0, 1.0, 2.0))
    assert rel_fro(b, a) < 1e-14


def test_rhs_potential_farfield_basis(ctx, oracle):
    geom = G.analytic_curve("trefoil", 40)
    k = 7.0
    dirs = np.array([[1.0, 0.0], [0.6, 0.8], [-0.28, 0.96]])
    for bases in (LIN, POL2, capi.wavepol_bases(1, 3.0)):
        mesh = ctx.mesh(geom, 12, 3, 0.15)
        nb = len(bases)
        for kind in (0, 1, 2):
            B = _np(ctx.rhs_basis(mesh, bases, kind, k, k, dirs, 2.0))
            ref = oracle.rhs_basis(geom, bases, 12, kind, k, k, dirs, 2.0)
            assert rel_fro(B, ref) < 1e-13, (nb, kind)
        kc = 7.0 - 0.4j
        B = _np(ctx.rhs_basis(mesh, bases, 2, kc, 7.0, dirs))
        assert rel_fro(B, oracle.rhs_basis(geom, bases, 12, 2, kc, 7.0, dirs)) < 1e-13
        rng = np.random.default_rng(nb)
        coef = rng.standard_normal((geom.n * nb, 2)) + 1j * rng.standard_normal((geom.n * nb, 2))
        c = torch.from_numpy(np.ascontiguousarray(coef.T)).to("cuda").t()
        pts = np.array([[2.5, 0.3], [-1.6, 1.2], [0.1, -2.2]])
        for kk in (k, 7.0 - 0.4j):
            u = _np(ctx.eval_single_layer_basis(mesh, bases, kk, c[:, 0].contiguous(),
                                                torch.from_numpy(pts).to("cuda")))
            assert rel_fro(u, oracle.eval_sl_basis(geom, bases, 12, kk, coef[:, 0], pts)) < 1e-13
        th = np.linspace(0, 2 * np.pi, 73)
        ff = _np(ctx.farfield_basis(mesh, bases, k, c, torch.from_numpy(th).to("cuda")))
        assert rel_fro(ff, oracle.farfield_basis(geom, bases, 12, k, coef, th)) < 1e-13


def test_basis_argument_errors(ctx):
    mesh = ctx.mesh(G.unit_circle(10), 4, 2, 0.15)
    with pytest.raises(capi.Bem2dError):
        ctx.assemble_basis(mesh, capi.wavepol_bases(4, 1.0), 1, 2.0)  # 10 bases > 8
    with pytest.raises(capi.Bem2dError):
        ctx.assemble_basis(mesh, [(9, 0, 0.0, 0.0)], 1, 2.0)  # unknown kind
    with pytest.raises(capi.Bem2dError):
        ctx.assemble_basis(mesh, LIN, 1, -2.0)  # k <= 0


# ---- the drop-in driver in the shape of examples/disk.cpp ------------------------------------
DRIVER = os.path.join(ROOT, "bem2d_amd", "host", "disk_driver")


def _run(*args):
    if not os.path.exists(DRIVER):
        pytest.fail("disk_driver is not built")
    out = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    d = json.loads(re.sub(r"-?nan", "NaN", out.stdout))
    res = {}
    for key, v in d.items():
        if key == "t":
            res[key] = np.array(v)
        elif isinstance(v, list):
            a = np.array(v, dtype=float).reshape(-1, 2)
            res[key] = a[:, 0] + 1j * a[:, 1]
        else:
            res[key] = v
    return res


@pytest.mark.parametrize("basis,ppw,k,L,N,curve,graded", [
    ("lin", 10, 5.0, 7, 30, "circle", 0),       # examples/disk.cpp's options at a small ppw
    ("pol2", 8, 4.0, 3, 16, "kite", 0),
    ("wave", 6, 6.0, 3, 16, "ellipse", 1),
    ("p0", 12, 5.0, 3, 16, "kite", 1),
])
def test_disk_driver_matches_oracle(oracle, ctx, basis, ppw, k, L, N, curve, graded):
    print("disk driver", basis, ppw, k, L, N, curve, graded, flush=True)
    d = _run(basis, ppw, k, L, N, curve, graded)
    args = {"circle": [0, 0, 0, 0], "kite": [0, 0, 0, 0], "ellipse": [2.0, 0, 0, 0]}[curve]
    if graded:
        geom = G.analytic_curve_ppw(curve, ppw, k, args, closed=True, L=2, sigma=0.15)
    else:
        n = d["n"]  # int(Length k ppw / 2 pi), as the driver computes it
        assert abs(n - oracle.curve_length(G.CURVES[curve], args) * k * ppw / 2.0 / np.pi) <= 1
        geom = G.analytic_curve(curve, n, args)
        # the mesh: element intervals at equal arc length (the oracle's partition)
        t = oracle.arclength_partition(G.CURVES[curve], args, n)
        assert np.abs(d["t"][:, 0] - t[:-1]).max() < 1e-14
    assert d["n"] == geom.n
    assert np.abs(d["t"][:, 0] - geom.param[:, 0]).max() < 1e-14
    bases = {"lin": LIN, "pol2": POL2, "wave": capi.wave_bases(k), "p0": capi.pol_bases(0)}[basis]
    A = oracle.assemble_basis(geom, bases, N, L, 0.15, oracle.K_COMBINED, k, k, 1.0, 2.0)
    b = oracle.rhs_basis(geom, bases, N, oracle.F_COMBINED_PLANE, k, k, [1.0, 0.0], 2.0)
    x = oracle.lu_solve(A, b)[:, 0]
    tol = _tol(oracle, geom, bases, N, L, 0.15, oracle.K_COMBINED, k, k, 1.0, 2.0)
    assert rel_fro(d["diag"], np.diag(A)) < tol
    assert rel_fro(d["col0"], A[:, 0]) < tol
    assert rel_fro(d["solution"], x) < 100 * tol
    th = 2 * np.pi * np.arange(360) / 360
    ff = oracle.farfield_basis(geom, bases, N, k, x, th)[:, 0]
    assert rel_fro(d["farfield"], ff) < 100 * tol
    if curve == "ellipse":  # (-1.5, 0) lies inside Ellipse(2) (x semi-axis 2): masked as well
        assert np.isnan(d["u_at_points"][0].real)
    else:
        u = np.exp(1j * k * -1.5) - oracle.eval_sl_basis(geom, bases, N, k, x,
                                                         np.array([[-1.5, 0.0]]))
        assert abs(d["u_at_points"][0] - u[0]) < 100 * tol * abs(u[0])
    # (0.1, 0.2) lies inside the obstacle: set_plotInterior gives NaN there (the reference's
    # EvalIncident sets the incident field to NaN + 0i, so the imaginary part stays finite)
    assert np.isnan(d["u_at_points"][1].real) and np.isfinite(d["u_at_points"][1].imag)
    if curve == "circle":
        ser = analytic_farfield(k, th)
        assert np.abs(d["farfield"] - ser).max() / np.abs(ser).max() < 1e-4


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_complex_k_tiled_kernel(ctx, oracle, kind):
    """Meshes with enough 8 x 8 tiles take the tiled complex-k kernel (both orders of a pair from
    one complex Hankel evaluation): P0 and LinBasis against sampled oracle columns, including the
    columns next to the diagonal (singular entries) and both ends of an open arc."""
    k = 9.0 - 1.1j
    for geom, bases in ((G.kite(720), None), (G.open_arc(600, 0.4), LIN)):
        mesh = ctx.mesh(geom, 6, 3, 0.15)
        nn = geom.n * (1 if bases is None else len(bases))
        cols = np.array([0, 1, 2, nn // 3, nn // 2 + 1, nn - 2, nn - 1])
        if bases is None:
            A = _np(ctx.assemble(mesh, kind, k, 9.0, 1.0, 2.0))
            ref = oracle.assemble(geom, 6, 3, 0.15, kind, k, 9.0, 1.0, 2.0, cols=cols)
        else:
            A = _np(ctx.assemble_basis(mesh, bases, kind, k, 9.0, 1.0, 2.0))
            ref = oracle.assemble_basis(geom, bases, 6, 3, 0.15, kind, k, 9.0, 1.0, 2.0,
                                        cols=cols)
        err = rel_fro(A[:, cols], ref)
        print("tiled complex k kind %d n %d nb %s: %.2e" % (kind, geom.n, bases and len(bases),
                                                          err), flush=True)
        assert err < 1e-12
        # the whole matrix against the per-entry kernel's entries through a real-k identity:
        # k = 9 - 1.1i is no special case, so check symmetry of SL (Gauss2D is symmetric)
        if kind == 1 and bases is None:
            n = geom.n
            d = np.abs(np.subtract.outer(np.arange(n), np.arange(n)))
            reg = (d > 1) & (d < n - 1)
            S = A - np.diag(np.diag(A))  # mass on the diagonal
            assert np.abs((S - S.T)[reg]).max() <= 1e-14 * np.abs(S).max()


def test_resonance_pipeline(ctx, oracle):
    """resonances/trapresonance.cpp on a graded triangle (Polygon(points, ppw, kmax, L, sigma),
    PolBasis(2)): Id + 2 CDL - 2i|k| SL with complex k, ChangeBasis, SingularValues.  The mass
    matrix against the oracle in full, the operator on sampled columns (the oracle's complex-k
    matrix of this size takes minutes), and the smallest singular value against numpy's SVD of the
    same operator in the mass matrix's orthonormal basis."""
    tri = [(0.0, 0.0), (1.0, 0.1), (0.3, 0.8)]
    geom = G.polygon_ppw(tri, 10, 8.0, 2, 0.15)
    assert geom.n == 3 * (150 + 4)
    bases = capi.pol_bases(2)
    nn = geom.n * 3
    mesh = ctx.mesh(geom, 5, 3, 0.15)
    M = ctx.assemble_basis(mesh, bases, 0, 1.0, 0.0, 1.0, 0.0)
    Mref = oracle.assemble_basis(geom, bases, 5, 3, 0.15, oracle.K_NONE, 1.0, 0.0, 1.0, 0.0)
    assert rel_fro(_np(M), Mref) < 1e-13
    cols = np.array([0, 1, 2, 3, 4, 5, 460, 700, nn - 3, nn - 2, nn - 1])
    for k in (5.2 - 0.6j, 7.6 - 1.8j):
        A = ctx.assemble_basis(mesh, bases, 4, k, abs(k), 1.0, 2.0)
        An = _np(A)
        ref = oracle.assemble_basis(geom, bases, 5, 3, 0.15, oracle.K_COMBINED, k, abs(k), 1.0,
                                    2.0, cols=cols)
        assert rel_fro(An[:, cols], ref) < 1e-12
        s = _np(ctx.singular_values(ctx.change_basis(A, M)))
        L = np.linalg.cholesky(Mref)
        KB = np.linalg.solve(L, np.linalg.solve(L, An.conj().T).conj().T)
        sref = np.linalg.svd(KB, compute_uv=False)
        print("k %s: smallest singular value %.12f (numpy %.12f)" % (k, s[-1], sref[-1]),
              flush=True)
        assert abs(s[-1] - sref[-1]) <= 1e-10 * sref[0]
