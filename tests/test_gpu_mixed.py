"""Mixed-precision solve (FP32 MFMA LU + FP64 refinement) on the GPU: solutions against the FP64
path and numpy, the backward-error target of 1e-12, multi-RHS, and the FP64 fallback."""
import numpy as np
import pytest
import torch

from bem2d_amd import capi
from bem2d_amd import geometry as G
from conftest import rel_fro

pytestmark = pytest.mark.gpu


def _dev(a):
    """Column-major complex128 device copy of a 2-D array."""
    a = np.asarray(a, dtype=np.complex128)
    t = torch.empty((a.shape[1], a.shape[0]), dtype=torch.complex128, device="cuda")
    t.copy_(torch.from_numpy(np.ascontiguousarray(a.T)))
    return t.t()


def _backward(A, X, B):
    return (np.abs(B - A @ X).max(axis=0) /
            (np.abs(A).sum(axis=1).max() * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max()


@pytest.mark.parametrize("n,nrhs", [(64, 1), (300, 1), (777, 3), (1000, 12)])
def test_random_diagonally_weighted(ctx, n, nrhs):
    print("case n=%d nrhs=%d" % (n, nrhs), flush=True)
    rng = np.random.default_rng(n)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A += 4 * np.eye(n)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    Ad = _dev(A)
    Bd = _dev(B)
    rep = ctx.solve_mixed(Ad, Bd, tol=1e-12)
    X = Bd.cpu().numpy()
    assert rep.fallback_fp64 == 0 and rep.converged == 1, (rep.iterations, rep.backward_error)
    assert rep.backward_error <= 1e-12
    assert _backward(A, X, B) <= 2e-12
    assert rel_fro(X, np.linalg.solve(A, B)) < 1e-10


@pytest.mark.parametrize("n", [64, 130, 1000])
def test_matrix_norm_reported(ctx, n):
    """|A|_inf of the report (row sums formed while the matrix is cast to FP32) against numpy,
    for sizes on and off the 64-block grid."""
    print("case n=%d" % n, flush=True)
    rng = np.random.default_rng(7 + n)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A += 4 * np.eye(n)
    A[n // 3, :] *= 5.0  # one dominant row
    B = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
    rep = ctx.solve_mixed(_dev(A), _dev(B), tol=1e-12)
    ref = np.abs(A).sum(axis=1).max()
    assert abs(rep.matrix_norm_inf - ref) <= 1e-13 * ref, (rep.matrix_norm_inf, ref)


@pytest.mark.parametrize("geom_name,n,k", [("circle", 200, 10.0), ("kite", 1000, 50.0),
                                           ("arc", 500, 60.0)])
def test_bem_matrices(ctx, geom_name, n, k):
    print("case", geom_name, n, k, flush=True)
    geom = {"circle": G.unit_circle, "kite": G.kite, "arc": G.open_arc}[geom_name](n)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    A = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0)
    ang = np.linspace(0, np.pi, 3)
    B = ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, np.stack([np.cos(ang), np.sin(ang)], 1), 2.0)
    X64 = ctx.solve_fp64(A, B.clone())
    Xm = B.clone()
    rep = ctx.solve_mixed(A, Xm, tol=1e-12)
    assert rep.fallback_fp64 == 0, rep.backward_error
    assert rep.backward_error <= 1e-12
    err = (torch.linalg.norm(Xm - X64) / torch.linalg.norm(X64)).item()
    assert err < 1e-10, err
    # factors stay resident: refine a second set of right-hand sides with the same factorisation
    lu = ctx.lu32(n)
    lu.factor(A)
    X2 = torch.empty_like(B)
    rep2 = lu.refine(A, B, X2, tol=1e-12)
    assert rep2.converged == 1
    assert (torch.linalg.norm(X2 - X64) / torch.linalg.norm(X64)).item() < 1e-10
    lu.close()


def test_zero_diagonal_entry_needs_no_fallback(ctx):
    """A zero on the diagonal of an otherwise diagonally weighted matrix: the static row order puts
    another row's large entry there, and the mixed path converges without the FP64 LU."""
    n = 130
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n)) + 4 * np.eye(n) + 0j
    A[0, 0] = 0.0
    B = rng.standard_normal((n, 2)) + 0j
    Bd = _dev(B)
    rep = ctx.solve_mixed(_dev(A), Bd, tol=1e-12)
    assert rep.fallback_fp64 == 0 and rep.converged == 1
    assert _backward(A, Bd.cpu().numpy(), B) <= 2e-12


def test_zero_pivot_falls_back_to_fp64(ctx):
    """A leading block [[1, 2], [2, 4 + 1e-10]]: column 1 claims row 1, column 0 takes row 0, and
    the FP32 cast of 4 + 1e-10 leaves an exactly zero second pivot, while the FP64 matrix is
    regular.  The one-shot solve falls back to the pivoted FP64 LU."""
    n = 130
    rng = np.random.default_rng(1)
    A = 4 * np.eye(n) + 0.01 * rng.standard_normal((n, n)) + 0j
    A[:2, :] = 0.0
    A[:, :2] = 0.0
    A[:2, :2] = [[1.0, 2.0], [2.0, 4.0 + 1e-10]]
    B = rng.standard_normal((n, 2)) + 0j
    Bd = _dev(B)
    rep = ctx.solve_mixed(_dev(A), Bd, tol=1e-12)
    assert rep.fallback_fp64 == 1
    assert _backward(A, Bd.cpu().numpy(), B) <= 1e-14


@pytest.mark.parametrize("n,nrhs", [(700, 64), (700, 360), (1024, 64)])
def test_many_right_hand_sides(ctx, n, nrhs):
    """The library-GEMM refinement path (rocBLAS ctrsm on the FP32 factors, zgemm residual) against
    the FP64 solve, with the far field checked at 1e-10 as on the large configuration."""
    print("case many rhs n=%d nrhs=%d" % (n, nrhs), flush=True)
    k = 40.0
    geom = G.open_arc(n)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    A = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0)
    th = 2 * np.pi * np.arange(nrhs) / nrhs
    B = ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, np.stack([np.cos(th), np.sin(th)], 1), 2.0)
    X64 = ctx.solve_fp64(A, B.clone())
    lu = ctx.lu32(n)
    lu.factor(A)
    X = torch.empty_like(B)
    rep = lu.refine(A, B, X, tol=1e-12)
    lu.close()
    assert rep.converged == 1 and rep.backward_error <= 1e-12
    assert (torch.linalg.norm(X - X64) / torch.linalg.norm(X64)).item() < 1e-10
    ang = torch.tensor(2 * np.pi * np.arange(90) / 90, device="cuda")
    f32 = ctx.farfield(mesh, k, X, ang)
    f64 = ctx.farfield(mesh, k, X64, ang)
    assert (torch.linalg.norm(f32 - f64) / torch.linalg.norm(f64)).item() < 1e-10
    mesh.close()


def test_many_right_hand_sides_after_refactorisation(ctx):
    """A second factorisation of another matrix in the same factor object, then a many-right-hand-
    side solve, uses the new factors; repeated solves with the same factors agree."""
    n, nrhs = 700, 64
    print("case many rhs refactorisation n=%d" % n, flush=True)
    rng = np.random.default_rng(21)
    lu = ctx.lu32(n)
    for t in range(2):
        A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
        A += (3 + t) * np.eye(n)
        B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        Ad = torch.tensor(np.ascontiguousarray(A.T), device="cuda").t()
        Bd = torch.tensor(np.ascontiguousarray(B.T), device="cuda").t()
        lu.factor(Ad)
        for _ in range(2):
            X = torch.empty_like(Bd)
            rep = lu.refine(Ad, Bd, X, tol=1e-12)
            assert rep.converged == 1 and rep.backward_error <= 1e-12
            assert rel_fro(X.cpu().numpy(), np.linalg.solve(A, B)) < 1e-10
    lu.close()


def test_paired_panels_large_trailing_matrix(ctx):
    """n = 5440 (85 tiles) with panels in eights while the trailing matrix has >= 75 tiles, in
    fours while it has >= 60 and in twos while it has >= 40: the first panels group in eights (one
    depth-512 update per eight panels), the next ones in fours, then twos, the last ones not at
    all -- against the FP64 solve."""
    n, k = 5440, 60.0
    print("case grouped panels n=%d" % n, flush=True)
    saved, saved_q = ctx.tune_lu_pair_tiles(), ctx.tune_lu_quad_tiles()
    saved_o = ctx.tune_lu_oct_tiles()
    ctx.tune_lu_pair_tiles(40)
    ctx.tune_lu_quad_tiles(60)
    ctx.tune_lu_oct_tiles(75)
    try:
        geom = G.kite(n)
        mesh = ctx.mesh(geom, 16, 3, 0.15)
        A = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0)
        B = ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, [[0.6, 0.8], [1.0, 0.0]], 2.0)
        X64 = ctx.solve_fp64(A, B.clone())
        Xm = B.clone()
        rep = ctx.solve_mixed(A, Xm, tol=1e-12)
        mesh.close()
    finally:
        ctx.tune_lu_pair_tiles(saved)
        ctx.tune_lu_quad_tiles(saved_q)
        ctx.tune_lu_oct_tiles(saved_o)
    assert rep.fallback_fp64 == 0 and rep.backward_error <= 1e-12
    assert (torch.linalg.norm(Xm - X64) / torch.linalg.norm(X64)).item() < 1e-10


@pytest.mark.parametrize("n", [700, 5440])
def test_factor_graph_replays_match_eager(ctx, n):
    """Repeated factorisations of one buffer replay a captured hipGraph (from the second call on);
    every replay, with new matrix contents and eager work and synchronisations in between, gives
    the same factors as an eager factorisation of the same matrix on a fresh buffer (compared
    through one correction x = U^-1 L^-1 b).  n = 5440 (85 tiles), with panels in eights from 75
    tiles, fours from 60 and twos from 40, covers every grouping in the graph."""
    print("case factor graph n=%d" % n, flush=True)
    saved, saved_q = ctx.tune_lu_pair_tiles(), ctx.tune_lu_quad_tiles()
    saved_o = ctx.tune_lu_oct_tiles()
    ctx.tune_lu_pair_tiles(40)
    ctx.tune_lu_quad_tiles(60)
    ctx.tune_lu_oct_tiles(75)
    try:
        _factor_graph_replays(ctx, n)
    finally:
        ctx.tune_lu_pair_tiles(saved)
        ctx.tune_lu_quad_tiles(saved_q)
        ctx.tune_lu_oct_tiles(saved_o)


def _factor_graph_replays(ctx, n):
    rng = np.random.default_rng(n + 1)
    Abuf = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    B = _dev(rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1)))
    lu = ctx.lu32(n)
    for t in range(4):
        A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
        A += 4 * np.eye(n)
        Ad = _dev(A)
        Abuf.copy_(Ad)
        lu.factor(Abuf)  # t = 0 eager, t = 1 capture + replay, then replays
        X = torch.zeros_like(B)
        lu.refine(Abuf, B, X, maxit=1, raise_on_fail=False)
        ref = ctx.lu32(n)
        ref.factor(Ad)  # first factorisation of its buffer: eager
        Xr = torch.zeros_like(B)
        ref.refine(Ad, B, Xr, maxit=1, raise_on_fail=False)
        torch.cuda.synchronize()
        ref.close()
        d = (torch.linalg.norm(X - Xr) / torch.linalg.norm(Xr)).item()
        print("  factorisation %d: |x - x_eager|/|x_eager| = %.2e" % (t, d), flush=True)
        assert d < 1e-12, (t, d)  # the same kernels on the same data: rounding-level at most
    assert lu.graph_replays() == 3
    lu.close()


def _kite_system(ctx, n=1000, k=50.0):
    geom = G.kite(n)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    A = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0)
    ang = np.array([0.3, 2.0])
    B = ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, np.stack([np.cos(ang), np.sin(ang)], 1), 2.0)
    torch.cuda.synchronize()
    mesh.close()
    return A.cpu().numpy(), B.cpu().numpy()


@pytest.mark.parametrize("order", ["reversed", "random"])
def test_row_permuted_bem_matrix(ctx, oracle, order):
    """Static pivoting: the rows of the kite N=1000 matrix reversed or randomly permuted (the
    natural order puts every column's largest entry on the diagonal, so a factorisation without
    interchanges in the given order would meet tiny pivots).  The mixed path recovers the order,
    converges without falling back, and matches the oracle's pivoted LU (zgetrf semantics) on the
    permuted system to 1e-10."""
    print("case row-permuted kite", order, flush=True)
    A, B = _kite_system(ctx)
    n = A.shape[0]
    p = np.arange(n)[::-1] if order == "reversed" else np.random.default_rng(11).permutation(n)
    Ap, Bp = A[p, :], B[p, :]
    xr = oracle.lu_solve(Ap, Bp)
    Bd = _dev(Bp)
    rep = ctx.solve_mixed(_dev(Ap), Bd, tol=1e-12)
    X = Bd.cpu().numpy()
    print("  iterations %d gmres %d growth %.3g backward %.2e" % (
        rep.iterations, rep.gmres_steps, rep.growth, rep.backward_error), flush=True)
    assert rep.fallback_fp64 == 0 and rep.converged == 1
    assert rep.backward_error <= 1e-12
    assert rel_fro(X, xr) < 1e-10


def test_static_pivoting_undoes_a_row_swap(ctx):
    """Every column of the BEM matrix claims a different row (its diagonal), so the static order
    of the matrix with two rows swapped is the natural order again: the factors, hence one
    correction from x = 0, are bit-identical to those of the unswapped matrix."""
    A, B = _kite_system(ctx, n=700, k=40.0)
    n = A.shape[0]
    p = np.arange(n)
    p[[5, 600]] = p[[600, 5]]
    X = []
    for M, R in ((A, B), (A[p, :], B[p, :])):
        lu = ctx.lu32(n)
        Md = _dev(M)
        lu.factor(Md)
        Xd = torch.zeros((R.shape[1], n), dtype=torch.complex128, device="cuda").t()
        lu.refine(Md, _dev(R), Xd, maxit=1, raise_on_fail=False)
        torch.cuda.synchronize()
        X.append(Xd.cpu().numpy())
        lu.close()
    assert np.array_equal(X[0], X[1])


def test_gmres_ir_on_ill_conditioned_matrix(ctx):
    """cond(A) = 1e7: one FP32 correction gains too little for plain refinement (it stalls), and
    GMRES-IR preconditioned by the FP32 factors reaches the FP64 target without the FP64 LU."""
    n = 300
    rng = np.random.default_rng(3)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q1 * np.logspace(0, -7, n)) @ Q2.conj().T
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    Bd = _dev(B)
    rep = ctx.solve_mixed(_dev(A), Bd, tol=1e-12)
    print("  iterations %d gmres %d growth %.3g backward %.2e fallback %d" % (
        rep.iterations, rep.gmres_steps, rep.growth, rep.backward_error, rep.fallback_fp64),
        flush=True)
    assert rep.converged == 1 and rep.backward_error <= 1e-12
    assert rep.fallback_fp64 == 0 and rep.gmres_steps > 0
    assert _backward(A, Bd.cpu().numpy(), B) <= 2e-12


def test_growth_routes_to_fp64(ctx):
    """Every column's largest entry in row 0 (a dominant first row): the static order keeps the
    first row on the diagonal once and leaves tiny pivots behind it, the factors' element growth
    exceeds 1e4, and the one-shot solve goes to the pivoted FP64 LU without refining."""
    n = 200
    rng = np.random.default_rng(5)
    A = 1e-6 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A[0, :] = 1.0 + rng.uniform(size=n)
    A[1:, 0] += 1.0
    B = rng.standard_normal((n, 1)) + 0j
    Bd = _dev(B)
    rep = ctx.solve_mixed(_dev(A), Bd, tol=1e-12)
    print("  growth %.3g fallback %d" % (rep.growth, rep.fallback_fp64), flush=True)
    assert rep.fallback_fp64 == 1 and rep.growth > 1e4
    # the FP64 LU is backward stable (the forward error of this cond ~1e8 matrix is not a test)
    assert _backward(A, Bd.cpu().numpy(), B) <= 1e-14
