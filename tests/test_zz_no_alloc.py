"""The mixed solve without allocation (bem2d_solve_mixed_static) against bem2d_solve_mixed: two calls
in a row on one GPU, the second smaller, each the same solution as the allocating entry point (the
same kernels and step sequence), within 1e-10 of the FP64 solve; sizes beyond the fixed workspace
are rejected."""
import numpy as np
import pytest
import torch

from bem2d_amd import capi
from bem2d_amd import geometry as G

pytestmark = pytest.mark.gpu


def _system(ctx, geom, k, nrhs):
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    A = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0)
    ang = np.linspace(0.2, 2.9, nrhs)
    B = ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, np.stack([np.cos(ang), np.sin(ang)], 1), 2.0)
    torch.cuda.synchronize()
    mesh.close()
    return A, B


def test_static_solve_matches_allocating_solve(ctx):
    for geom, k, nrhs in ((G.kite(1000), 50.0, 3), (G.unit_circle(300), 10.0, 1)):
        n = geom.n
        print("case static mixed solve n=%d nrhs=%d" % (n, nrhs), flush=True)
        A, B = _system(ctx, geom, k, nrhs)
        Xs = B.clone()
        rc, rep = ctx.solve_mixed_static(A, Xs, tol=1e-12)
        Xm = B.clone()
        rep_m = ctx.solve_mixed(A, Xm, tol=1e-12)
        X64 = ctx.solve_fp64(A, B.clone())
        torch.cuda.synchronize()
        d = (torch.linalg.norm(Xs - Xm) / torch.linalg.norm(Xm)).item()
        e = (torch.linalg.norm(Xs - X64) / torch.linalg.norm(X64)).item()
        print("  steps %d/%d backward %.2e, |x_static - x_mixed| %.1e (bit-identical: %s), "
              "|x_static - x_fp64| %.1e" % (rep.iterations, rep_m.iterations, rep.backward_error,
                                           d, bool(torch.equal(Xs, Xm)), e), flush=True)
        assert rc == capi.OK and rep.converged == 1 and rep.backward_error <= 1e-12
        assert rep_m.fallback_fp64 == 0 and rep.iterations == rep_m.iterations
        assert d <= 1e-15 and e <= 1e-10


def test_static_solve_rejects_sizes_beyond_its_workspace(ctx):
    A = torch.eye(1100, dtype=torch.complex128, device="cuda").t()
    B = torch.ones((1, 1100), dtype=torch.complex128, device="cuda").t()
    rc, _ = ctx.solve_mixed_static(A, B, raise_on_fail=False)
    assert rc == capi.ERR_ARG
    A = torch.eye(64, dtype=torch.complex128, device="cuda").t()
    B = torch.ones((9, 64), dtype=torch.complex128, device="cuda").t()
    rc, _ = ctx.solve_mixed_static(A, B, raise_on_fail=False)
    assert rc == capi.ERR_ARG
