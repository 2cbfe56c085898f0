"""Complex wavenumbers on the GPU through the C-ABI: the device complex Hankel routine against
mpmath, complex-k matrices against entries evaluated in mpmath and against the oracle
(SingleLayer / DoubleLayer / ConjDoubleLayer / CombinedSingleConjDouble with freqtype k.im != 0,
lib/bem2d_kernel.cpp:15-112), complex-k right-hand sides and potentials, and a complex-k
combined-field solve."""
import numpy as np
import pytest

from bem2d_amd import capi
from bem2d_amd import geometry as G
from conftest import golden

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative Frobenius, FP64 against the oracle


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_device_hankel_complex_matches_mpmath(ctx):
    import torch
    f = golden("hankel_complex_mpmath.npz")
    z, v = f["z"], f["h0h1"]
    phis = np.round(np.angle(z), 12)
    worst = 0.0
    for phi in np.unique(phis):
        sel = phis == phi
        zt = torch.tensor(z[sel], dtype=torch.complex128, device="cuda")
        out = ctx.hankel01_complex(float(phi), zt).cpu().numpy()
        for c in (0, 2):
            ref = v[sel, c] + 1j * v[sel, c + 1]
            got = out[:, c] + 1j * out[:, c + 1]
            worst = max(worst, (np.abs(got - ref) / np.abs(ref)).max())
    print("device complex Hankel vs mpmath: %.2e" % worst, flush=True)
    assert worst < 1e-14


@pytest.mark.parametrize("name", ["circle4_ck", "triangle6_ck", "kite5_ck"])
def test_complex_entries_match_mpmath(ctx, name):
    f = golden("entries_complex_mpmath.npz")
    geom = G.Geometry(f[name + "_type"], f[name + "_param"], f[name + "_index"],
                      f[name + "_next"], f[name + "_prev"])
    N, L, sigma = f[name + "_quad"]
    kr, ki, eta = f[name + "_kc"]
    mesh = ctx.mesh(geom, int(N), int(L), float(sigma))
    for key in f.files:
        if not key.startswith(name + "_kind"):
            continue
        kind = int(key[len(name) + 5:])
        A = ctx.assemble(mesh, kind, complex(kr, ki), eta, 0.0, 1.0).cpu().numpy()
        err = _rel(A, f[key])
        print(name, kind, "%.2e" % err, flush=True)
        assert err < 1e-13, (name, kind, err)
    mesh.close()


CASES = [
    # name, geometry, k, kinds
    ("circle48", lambda: G.unit_circle(48), 6.0 - 1.5j, (capi.KERNEL_SL, capi.KERNEL_COMBINED)),
    ("kite40", lambda: G.kite(40), 7.5 - 0.5j, (capi.KERNEL_DL, capi.KERNEL_CDL)),
    ("square36", lambda: G.polygon([(0, 0), (1, 0), (1, 1), (0, 1)], 9), 4.0 + 2.0j,
     (capi.KERNEL_COMBINED,)),
    ("arc32open", lambda: G.open_arc(32), 5.0 - 2.0j, (capi.KERNEL_SL, capi.KERNEL_DL)),
    ("circle24_farz", lambda: G.unit_circle(24), 40.0 - 3.0j, (capi.KERNEL_COMBINED,)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_complex_assembly_matches_oracle(ctx, oracle, case):
    name, build, k, kinds = case
    geom = build()
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    for kind in kinds:
        ms, ks = (1.0, 2.0) if kind == capi.KERNEL_COMBINED else (0.0, 1.0)
        A = ctx.assemble(mesh, kind, k, abs(k), ms, ks).cpu().numpy()
        R = oracle.assemble(geom, 16, 3, 0.15, kind, k, abs(k), ms, ks)
        err = _rel(A, R)
        print(name, "k=%s" % k, "kind", kind, "%.2e" % err, flush=True)
        assert err < TOL, (name, kind, err)
    mesh.close()


def test_complex_rhs_and_potential_match_oracle(ctx, oracle):
    import torch
    geom = G.kite(40)
    k, eta = 7.0 - 1.0j, 7.0
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    dirs = np.array([[1.0, 0.0], [np.cos(2.0), np.sin(2.0)]])
    for kind in (capi.FIELD_PLANE, capi.FIELD_NORMAL_PLANE, capi.FIELD_COMBINED_PLANE):
        B = ctx.rhs(mesh, kind, k, eta, dirs, 2.0).cpu().numpy()
        R = oracle.rhs(geom, 16, kind, k, eta, dirs, 2.0)
        assert _rel(B, R) < 1e-13, kind
    rng = np.random.default_rng(3)
    coef = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    pts = np.stack([3 * np.cos(np.linspace(0, 6, 17)), 2.5 * np.sin(np.linspace(0, 6, 17))], 1)
    u = ctx.eval_single_layer(mesh, k, torch.tensor(coef, device="cuda"),
                              torch.tensor(pts, device="cuda")).cpu().numpy()
    ur = oracle.eval_sl(geom, 16, k, coef, pts)
    assert _rel(u, ur) < TOL
    mesh.close()


def test_complex_k_solve_matches_oracle(ctx, oracle):
    """A = M + 2(CDL - i eta SL) with complex k (the resonance scans' combined operator,
    resonances/trapresonance.cpp:55-68), solved by the plain FP64 path."""
    geom = G.unit_circle(40)
    k, eta = 6.0 - 0.7j, abs(6.0 - 0.7j)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    A = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, eta, 1.0, 2.0)
    B = ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, eta, [[1.0, 0.0]], 2.0)
    X = ctx.solve_fp64(A, B.clone()).cpu().numpy()
    Ar = oracle.assemble(geom, 16, 3, 0.15, oracle.K_COMBINED, k, eta, 1.0, 2.0)
    br = oracle.rhs(geom, 16, oracle.F_COMBINED_PLANE, k, eta, [1.0, 0.0], 2.0)
    xr = oracle.lu_solve(Ar, br)
    assert _rel(X, xr) < TOL
    mesh.close()


def test_complex_entry_points_edges(ctx):
    import torch
    geom = G.unit_circle(16)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    # k.im == 0 through the complex entry point is the real-k path, bit for bit
    A1 = ctx.assemble(mesh, capi.KERNEL_COMBINED, 5.0, 5.0, 1.0, 2.0).cpu().numpy()
    out = torch.empty((16, 16), dtype=torch.complex128, device="cuda").t()
    capi.check(capi.lib().bem2d_assemble_complex(ctx.h, mesh.h, capi.KERNEL_COMBINED, 5.0, 0.0,
                                                 5.0, 1.0, 2.0, capi._t_ptr(out), 16))
    assert np.array_equal(A1, out.cpu().numpy())
    # Re k <= 0 is outside the supported domain
    with pytest.raises(capi.Bem2dError) as e:
        ctx.assemble(mesh, capi.KERNEL_SL, -1.0 + 2.0j, 0.0, 0.0, 1.0)
    assert e.value.code == capi.ERR_UNSUPPORTED
    mesh.close()
