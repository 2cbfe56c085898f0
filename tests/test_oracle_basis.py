"""The oracle's basis-function restatement (ConstBasis, PolBasis, LinBasis, WaveBasis, WavePolBasis
in the Galerkin entries, mass matrix, right-hand sides, potential and far field) and its arc-length
parameterisation, pinned on the CPU:

* entries against the mpmath fixtures (entries_basis_mpmath.npz, 30 digits);
* identities the basis definitions fix: LinBasis(+1) + LinBasis(-1) = 1, PolBasis(1) = LinBasis(+1),
  PolBasis(0) = ConstBasis = the P0 path, WavePolBasis(0, d, k) = WaveBasis(d, k), and the
  matching sums of the right-hand sides, potential and far field;
* the conjugated test function, through closed forms on flat elements (wave-basis mass matrix,
  a plane wave whose phase a wave basis cancels);
* test / trial orientation: with real bases ConjDoubleLayer is the transpose of DoubleLayer on the
  regular pairs (K_CDL(x, y) = K_DL(y, x));
* the arc-length partition against closed forms (circle, ellipse through scipy's incomplete elliptic
  integral) and the host C-ABI mesher against the oracle.
"""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.optimize as so
import scipy.special as ss

from bem2d_amd import capi
from bem2d_amd import geometry as G
from conftest import golden, rel_fro

LIN = [(2, 1, 0.0, 0.0), (2, -1, 0.0, 0.0)]


def _geom(f, name):
    return G.Geometry(f[name + "_type"], f[name + "_param"], f[name + "_index"], f[name + "_next"],
                      f[name + "_prev"])


@pytest.mark.parametrize("name", ["circle4_lin", "triangle6_pol2", "arc4open_wavepol1",
                                  "circle3_wave_ck"])
def test_basis_entries_against_mpmath(oracle, name):
    f = golden("entries_basis_mpmath.npz")
    geom = _geom(f, name)
    N, L, sigma = f[name + "_quad"]
    kr, ki, eta = f[name + "_kc"]
    bases = [(int(b[0]), int(b[1]), b[2], complex(b[3], b[4])) for b in f[name + "_bases"]]
    k = complex(kr, ki) if ki != 0 else kr
    kinds = sorted(int(key.split("kind")[1]) for key in f.files if key.startswith(name + "_kind"))
    assert kinds
    for kind in kinds:
        ref = f["%s_kind%d" % (name, kind)]
        if kind == 0:
            got = oracle.assemble_basis(geom, bases, int(N), int(L), sigma, oracle.K_NONE, 1.0,
                                        0.0, 1.0, 0.0)
        else:
            got = oracle.assemble_basis(geom, bases, int(N), int(L), sigma, kind, k, eta, 0.0, 1.0)
        err = rel_fro(got, ref)
        print(name, kind, "%.2e" % err)
        assert err < 1e-13, (name, kind, err)


def test_lin_and_pol_identities(oracle):
    geom = G.kite(24)
    p0 = oracle.assemble(geom, 8, 3, 0.15, oracle.K_COMBINED, 6.0, 6.0, 1.0, 2.0)
    lin = oracle.assemble_basis(geom, LIN, 8, 3, 0.15, oracle.K_COMBINED, 6.0, 6.0, 1.0, 2.0)
    # t + (1 - t) = 1: the four blocks of each element pair add up to the P0 entry
    s = lin[0::2, 0::2] + lin[0::2, 1::2] + lin[1::2, 0::2] + lin[1::2, 1::2]
    assert rel_fro(s, p0) < 1e-14
    pol = oracle.assemble_basis(geom, [(1, 0, 0, 0), (1, 1, 0, 0)], 8, 3, 0.15, oracle.K_COMBINED,
                                6.0, 6.0, 1.0, 2.0)
    assert np.abs(pol[0::2, 0::2] - p0).max() == 0.0          # PolBasis(0) = the P0 path
    assert np.abs(pol[1::2, 1::2] - lin[0::2, 0::2]).max() == 0.0  # PolBasis(1) = LinBasis(+1)
    k = 3.0
    wave = oracle.assemble_basis(geom, [(3, 0, 1.0, k), (3, 0, -1.0, k)], 8, 3, 0.15,
                                 oracle.K_SL, 6.0, 0.0, 0.0, 1.0)
    wp = oracle.assemble_basis(geom, [(4, 0, 1.0, k), (4, 0, -1.0, k)], 8, 3, 0.15, oracle.K_SL,
                               6.0, 0.0, 0.0, 1.0)
    assert rel_fro(wp, wave) < 1e-15


def test_wave_mass_closed_form(oracle):
    """Mass matrix of WaveBasis(+-1, kb) on a flat element of length h: h int_0^1
    conj(e^{i kb d_a t}) e^{i kb d_b t} dt = h for d_a = d_b and h (e^{-2i kb d_a} - 1)/(-2i kb d_a)
    otherwise (the conjugate is what makes the diagonal real)."""
    geom = G.line((0.0, 0.0), (0.6, 0.8), 3)  # h = 1/3
    kb = 2.5 - 0.4j
    bases = [(3, 0, 1.0, kb), (3, 0, -1.0, kb)]
    M = oracle.assemble_basis(geom, bases, 30, 3, 0.15, oracle.K_NONE, 1.0, 0.0, 1.0, 0.0)
    h = 1.0 / 3
    for e in range(3):
        blk = M[2 * e:2 * e + 2, 2 * e:2 * e + 2]
        for a, da in enumerate((1.0, -1.0)):
            for b, db in enumerate((1.0, -1.0)):
                z = 1j * (kb * db - np.conj(kb) * da)   # conj(e^{i kb da t}) e^{i kb db t} = e^{z t}
                ref = h * (np.expm1(z) / z if abs(z) > 0 else 1.0)
                assert abs(blk[a, b] - ref) < 1e-14, (e, a, b, blk[a, b], ref)
    assert np.abs(M[0:2, 2:]).max() == 0.0


def test_rhs_conjugate_closed_form(oracle):
    """A plane wave along a flat element (x = p1 + t D, d = D/h) against WaveBasis(+1, k h): the
    conjugated basis cancels the phase, so <f, phi> = h e^{ik d.p1} exactly for any N."""
    k = 7.0
    geom = G.line((0.2, -0.1), (0.2 + 0.6, -0.1 + 0.8), 2)  # h = 0.5, d = (0.6, 0.8)
    h = 0.5
    b = oracle.rhs_basis(geom, [(3, 0, 1.0, k * h)], 5, oracle.F_PLANE, k, 0.0, [0.6, 0.8])
    for e in range(2):
        p1 = np.array([0.2, -0.1]) + e * h * np.array([0.6, 0.8])
        assert abs(b[e, 0] - h * np.exp(1j * k * (0.6 * p1[0] + 0.8 * p1[1]))) < 1e-14


def test_rhs_potential_farfield_linearity(oracle):
    geom = G.curve_mesh("trefoil", 30)
    k = 5.0
    r0 = oracle.rhs(geom, 12, oracle.F_COMBINED_PLANE, k, k, [[0.6, 0.8]])
    r1 = oracle.rhs_basis(geom, LIN, 12, oracle.F_COMBINED_PLANE, k, k, [[0.6, 0.8]])
    assert rel_fro(r1[0::2] + r1[1::2], r0) < 1e-14
    c = np.random.default_rng(3).standard_normal(30) + 1j * np.random.default_rng(4).standard_normal(30)
    cl = np.repeat(c, 2)  # the same coefficient on t and 1 - t: the P0 density
    pts = np.array([[2.0, 0.5], [-1.7, -1.1]])
    assert rel_fro(oracle.eval_sl_basis(geom, LIN, 12, k, cl, pts),
                   oracle.eval_sl(geom, 12, k, c, pts)) < 1e-14
    th = np.linspace(0, 2 * np.pi, 17)
    assert rel_fro(oracle.farfield_basis(geom, LIN, 12, k, cl, th),
                   oracle.farfield(geom, 12, k, c, th)) < 1e-14
    # complex k: the complex branches of the kernel in the potential
    kc = 5.0 - 0.3j
    assert rel_fro(oracle.eval_sl_basis(geom, LIN, 12, kc, cl, pts),
                   oracle.eval_sl(geom, 12, kc, c, pts)) < 1e-14


def test_cdl_is_dl_transposed_on_regular_pairs(oracle):
    """K_CDL(x, y) = K_DL(y, x), so with real bases A_CDL[(i,a),(j,b)] = A_DL[(j,b),(i,a)] on the
    element pairs that use Gauss2D (symmetric in the two roles); a transposed basis index fails."""
    n = 20
    geom = G.kite(n)
    bases = [(1, 0, 0, 0), (1, 1, 0, 0), (1, 2, 0, 0)]
    cdl = oracle.assemble_basis(geom, bases, 10, 3, 0.15, oracle.K_CDL, 4.0, 0.0, 0.0, 1.0)
    dl = oracle.assemble_basis(geom, bases, 10, 3, 0.15, oracle.K_DL, 4.0, 0.0, 0.0, 1.0)
    el = np.arange(n * 3) // 3
    d = (el[:, None] - el[None, :]) % n
    regular = (d > 1) & (d < n - 1)
    diff = np.abs(cdl - dl.T)[regular].max()
    assert diff < 1e-15 * np.abs(cdl).max() * 100
    # the same comparison with the basis index transposed inside each block is far off
    perm = (el * 3 + (2 - np.arange(n * 3) % 3))
    off = np.abs(cdl - dl.T[perm][:, perm])[regular].max()
    assert off > 1e-3 * np.abs(cdl).max()


def test_curve_length_closed_forms(oracle):
    assert abs(oracle.curve_length(G.ELEM_CIRCLE, [0, 0, 0, 0]) - 2 * np.pi) < 4e-15
    for rho in (1.5, 2.0, 3.0):
        a, b = rho, 1 / rho
        ref = 4 * a * ss.ellipe(1 - b * b / (a * a))
        assert abs(oracle.curve_length(G.ELEM_ELLIPSE, [rho, 0, 0, 0]) - ref) < 1e-14 * ref
    # kite and trefoil against scipy's adaptive quadrature of |Gamma'|
    for ty in (G.ELEM_KITE, G.ELEM_TREFOIL):
        def speed(t, ty=ty):
            return np.linalg.norm(oracle.elem_eval(G.curve_mesh(ty, 1), 0, t)[2:4])
        ref = si.quad(speed, 0, 1, epsabs=0, epsrel=2e-14, limit=400)[0]
        assert abs(oracle.curve_length(ty, [0, 0, 0, 0]) - ref) < 1e-13 * ref


def test_arclength_partition_closed_forms(oracle):
    t = oracle.arclength_partition(G.ELEM_CIRCLE, [0, 0, 0, 0], 12)
    assert np.abs(t - np.arange(13) / 12).max() < 1e-15
    # ellipse x = rho cos(2 pi t), y = sin(2 pi t)/rho: s(t) = rho E(2 pi t | 1 - 1/rho^4) on the
    # first quarter (the parameterisation by phi = 2 pi t - pi/2 of the standard form)
    rho = 2.0
    m = 1 - 1 / rho ** 4
    L = 4 * rho * ss.ellipe(m)

    def s_of(t):  # arc length from t = 0, by the incomplete integral E(phi | m) in x = a sin(phi)
        phi = 2 * np.pi * t
        return rho * (ss.ellipeinc(phi - np.pi / 2, m) + ss.ellipe(m))

    n = 16
    t = oracle.arclength_partition(G.ELEM_ELLIPSE, [rho, 0, 0, 0], n)
    for i in range(1, n):
        ref = so.brentq(lambda x: s_of(x) - i * L / n, 0, 1, xtol=1e-16, rtol=1e-15)
        assert abs(t[i] - ref) < 2e-15, (i, t[i], ref)


def test_capi_mesher_matches_oracle(oracle):
    """The library's host mesher (bem2d_curve_length, bem2d_arclength_params) against the oracle."""
    for ty, args in ((G.ELEM_KITE, [0, 0, 0, 0]), (G.ELEM_ELLIPSE, [2.0, 0, 0, 0]),
                     (G.ELEM_INV_ELLIPSE, [0.4, 0, 0, 0]), (G.ELEM_TREFOIL, [0, 0, 0, 0]),
                     (G.ELEM_ELLIPSE_ARC, [1.3, 0.7, 0.3, 5.5])):
        L = capi.curve_length(ty, args)
        assert abs(L - oracle.curve_length(ty, args)) < 1e-14 * L
        n = 37
        s = np.arange(1, n) * L / n
        t = capi.arclength_params(ty, args, s)
        ref = oracle.arclength_partition(ty, args, n)[1:-1]
        assert np.abs(t - ref).max() < 1e-14, ty
        assert np.all(np.diff(t) > 0)
    with pytest.raises(capi.Bem2dError):
        capi.arclength_params(G.ELEM_KITE, [0, 0, 0, 0], [0.5, 0.2])  # decreasing
    with pytest.raises(capi.Bem2dError):
        capi.curve_length(G.ELEM_FLAT, [0, 0, 0, 0])


def test_analytic_curve_builders(oracle):
    """AnalyticCurve(n, curve) and AnalyticCurve(ppw, k, curve, closed, L, sigma) partitions."""
    g = G.analytic_curve("kite", 40)
    ref = oracle.arclength_partition(G.ELEM_KITE, [0, 0, 0, 0], 40)
    assert np.abs(g.param[:, 0] - ref[:-1]).max() < 1e-14
    assert np.abs(g.param[:, 1] - ref[1:]).max() < 1e-14 and g.param[-1, 1] == 1.0
    assert list(g.next[:3]) == [1, 2, 3] and g.prev[0] == 39
    # ppw form: c = max(ceil(ppw k L / 2 pi), 10) pieces, the first and last split into L + 1
    # graded pieces of ratio sigma
    L = oracle.curve_length(G.ELEM_CIRCLE, [0, 0, 0, 0])
    go = G.analytic_curve_ppw("circle", 6, 3.0, closed=False, L=2, sigma=0.15)
    c = int(np.ceil(6 * 3.0 * L / 2.0 / np.pi))
    assert go.n == c + 4
    lengths = (go.param[:, 1] - go.param[:, 0]) * 2 * np.pi
    h = L / c
    assert np.allclose(lengths[:3], [0.15 ** 2 * h, (0.15 - 0.15 ** 2) * h, (1 - 0.15) * h],
                       rtol=1e-12)
    assert np.allclose(lengths[-3:], [(1 - 0.15) * h, (0.15 - 0.15 ** 2) * h, 0.15 ** 2 * h],
                       rtol=1e-12)
    assert go.prev[0] == -1 and go.next[-1] == -1
