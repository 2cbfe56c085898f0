"""Writes the mpmath golden vectors under tests/golden/ (run in a container that has mpmath).

    python tests/golden/make_fixtures.py

Nothing here reads or runs the reference.  The fixture files produced:

hankel_mpmath.npz   J0, Y0, J1, Y1 at 50 significant digits, rounded to float64, at log-spaced
                    x in [1e-8, 5e4] plus dense clusters around the algorithm switch points of
                    the oracle (2, 25) and of the device code, and around the first zeros of
                    J0/Y0/J1/Y1.
gauss_mpmath.npz    Gauss-Legendre nodes and weights on [0, 1] by Newton's method on P_n at 40
                    digits.
hankel_complex_mpmath.npz
                    H0^(1), H1^(1) of complex z, |arg z| <= 0.95 pi/2, 1e-6 <= |z| <= 2e3,
                    -650 < Im z < 120 (plus clusters at the device code's switch radii 2 and 20),
                    30 significant digits:
                    mpmath's J + iY loses ~2 Im z / ln 10 digits to cancellation, so the working
                    precision is raised by that much.
entries_complex_mpmath.npz
                    as entries_mpmath.npz for complex wavenumbers (the kernels' k.im != 0 branch).
entries_basis_mpmath.npz
                    as entries_mpmath.npz with several basis functions per element (PolBasis,
                    LinBasis, WaveBasis, WavePolBasis; row = element * nb + basis): the entries
                    carry conj(phi_a(test parameter)) phi_b(trial parameter), the mass matrix
                    conj(phi_a) phi_b, evaluated in mpmath at 30 digits.
entries_mpmath.npz  Galerkin matrices of tiny meshes evaluated entirely in mpmath (30 digits):
                    the reference's quadrature rules are rebuilt from the 40-digit Gauss nodes,
                    geometry, kernels and Hankel functions are evaluated in mpmath.  These pin the
                    double-precision evaluation chain of the oracle (and of the HIP path) at the
                    1e-14 level; the kernel conventions themselves are pinned separately by the
                    analytic circle identities in the tests.
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def gauss_mp(n, dps=40):
    with mp.workdps(dps):
        if n == 1:
            return [mp.mpf("0.5")], [mp.mpf(1)]
        xs, ws = [], []
        for i in range(1, n + 1):
            x = mp.cos(mp.pi * (i - mp.mpf("0.25")) / (n + mp.mpf("0.5")))
            for _ in range(100):
                p0, p1 = mp.mpf(1), x
                for k in range(2, n + 1):
                    p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
                dp = n * (x * p1 - p0) / (x * x - 1)
                dx = p1 / dp
                x -= dx
                if abs(dx) < mp.mpf(10) ** (-dps + 3):
                    break
            p0, p1 = mp.mpf(1), x
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            dp = n * (x * p1 - p0) / (x * x - 1)
            w = 2 / ((1 - x * x) * dp * dp)
            xs.append((1 + x) / 2)
            ws.append(w / 2)
        order = sorted(range(n), key=lambda i: xs[i])
        return [xs[i] for i in order], [ws[i] for i in order]


def rule_mp(which, N, L, sigma):
    """The reference's tensor / graded rules (lib/bem2d_quadpoints.cpp:166-318) in mpmath."""
    xt, wt = gauss_mp(N)
    gx = [xt[i] for i in range(N) for j in range(N)]
    gy = [xt[j] for i in range(N) for j in range(N)]
    gw = [wt[i] * wt[j] for i in range(N) for j in range(N)]
    X, Y, W = [], [], []

    def box(a, b, c, d):
        for i in range(len(gx)):
            X.append(a + (b - a) * gx[i])
            Y.append(c + (d - c) * gy[i])
            W.append((b - a) * (d - c) * gw[i])

    s = mp.mpf(sigma)
    if which == 0:
        return gx, gy, gw
    if which == 1:
        xl = [s ** (L - i) for i in range(L)] + [mp.mpf(1)]
        for i in range(L):
            for j in range(L):
                box(xl[i], xl[i + 1], xl[j], xl[j + 1])
    elif which == 2:
        a, b = s, mp.mpf(1)
        for _ in range(L):
            box(0, a, 1 - b, 1 - a)
            box(a, b, 1 - b, 1 - a)
            box(a, b, 1 - a, 1)
            a *= s
            b *= s
    elif which == 3:
        a, b = mp.mpf(1), s
        for _ in range(L):
            box(1 - a, 1 - b, 0, b)
            box(1 - a, 1 - b, b, a)
            box(1 - b, 1, b, a)
            a *= s
            b *= s
    return X, Y, W


# ---- geometry in mpmath (same parameterisations as the element types of the C-ABI) -------------

def curve(ty, c, t):
    tp = 2 * mp.pi
    if ty == 1:
        return (mp.cos(tp * t), mp.sin(tp * t)), (-tp * mp.sin(tp * t), tp * mp.cos(tp * t))
    if ty == 3:
        M = (mp.cos(tp * t) + mp.mpf("0.65") * mp.cos(2 * tp * t) - mp.mpf("0.65"),
             mp.mpf("1.5") * mp.sin(tp * t))
        D = (-tp * mp.sin(tp * t) - mp.mpf("0.65") * 2 * tp * mp.sin(2 * tp * t),
             mp.mpf("1.5") * tp * mp.cos(tp * t))
        return M, D
    if ty == 5:
        a, b, t1, t2 = [mp.mpf(v) for v in c[:4]]
        th = t1 + t * (t2 - t1)
        return (a * mp.cos(th), b * mp.sin(th)), (-a * (t2 - t1) * mp.sin(th),
                                                  b * (t2 - t1) * mp.cos(th))
    raise ValueError(ty)


def elem(ty, p, t):
    if ty == 0:
        p1x, p1y, dx, dy = [mp.mpf(v) for v in p[:4]]
        return (p1x + t * dx, p1y + t * dy), (dx, dy)
    ts, te = mp.mpf(p[0]), mp.mpf(p[1])
    M, D = curve(ty, p[2:], ts + t * (te - ts))
    return M, ((te - ts) * D[0], (te - ts) * D[1])


def kernel(kind, k, eta, x, y, n1, n2):
    i = mp.mpc(0, 1)
    if kind == 5:
        return mp.mpc(1)
    w = (x[0] - y[0], x[1] - y[1])
    d = mp.sqrt(w[0] ** 2 + w[1] ** 2)
    z = k * d
    if kind == 1:
        return i / 4 * mp.hankel1(0, z)
    h1 = mp.hankel1(1, z)
    if kind == 2:
        return -i * k / 4 * h1 / d * (-(n2[0] * w[0] + n2[1] * w[1]))
    cdl = -i * k / 4 * h1 / d * (n1[0] * w[0] + n1[1] * w[1])
    if kind == 3:
        return cdl
    return cdl - i * eta * (i / 4 * mp.hankel1(0, z))


def normal(D):
    l = mp.sqrt(D[0] ** 2 + D[1] ** 2)
    return (D[1] / l, -D[0] / l), l


def matrix_mp(types, params, index, nxt, prv, N, L, sigma, kind, k, eta):
    n = len(types)
    k = mp.mpc(k.real, k.imag) if isinstance(k, complex) else mp.mpf(k)
    eta = mp.mpf(eta)
    rules = [rule_mp(w, N, L, sigma) for w in range(4)]
    A = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            acc = mp.mpc(0)
            if index[i] == index[j]:
                X, Y, W = rules[1]
                for q in range(len(X)):
                    for (s, t) in ((X[q], (1 - Y[q]) * X[q]), (1 - X[q], 1 - X[q] + Y[q] * X[q])):
                        Mx, Dx = elem(types[i], params[i], s)
                        My, Dy = elem(types[j], params[j], t)
                        n1, s1 = normal(Dx)
                        n2, s2 = normal(Dy)
                        acc += X[q] * kernel(kind, k, eta, Mx, My, n1, n2) * s1 * s2 * W[q]
            else:
                r = 0
                if nxt[i] == index[j]:
                    r = 3
                elif prv[i] == index[j]:
                    r = 2
                X, Y, W = rules[r]
                for q in range(len(X)):
                    Mx, Dx = elem(types[i], params[i], X[q])
                    My, Dy = elem(types[j], params[j], Y[q])
                    n1, s1 = normal(Dx)
                    n2, s2 = normal(Dy)
                    acc += kernel(kind, k, eta, Mx, My, n1, n2) * s1 * s2 * W[q]
            A[i, j] = complex(acc)
    return A


def ring(n):
    idx = np.arange(n, dtype=np.int64)
    return idx, (idx + 1) % n, (idx - 1) % n


def curve_mesh(ty, n, c=(0, 0, 0, 0), closed=True):
    types = np.full(n, ty, dtype=np.int32)
    params = np.zeros((n, 8))
    for e in range(n):
        params[e, 0] = e / n
        params[e, 1] = (e + 1) / n
        params[e, 2:6] = c
    idx, nx, pv = ring(n)
    if not closed:
        nx[-1] = -1
        pv[0] = -1
    return types, params, idx, nx, pv


def polygon_mesh(verts, per_side):
    types, params = [], []
    m = len(verts)
    for s in range(m):
        p0 = np.array(verts[s], dtype=float)
        p1 = np.array(verts[(s + 1) % m], dtype=float)
        d = p1 - p0
        for j in range(per_side):
            a = p0 + (j / per_side) * d
            b = p0 + ((j + 1) / per_side) * d
            types.append(0)
            params.append([a[0], a[1], b[0] - a[0], b[1] - a[1], 0, 0, 0, 0])
    n = len(types)
    idx, nx, pv = ring(n)
    return np.array(types, dtype=np.int32), np.array(params), idx, nx, pv


CASES = [
    # name, mesh builder, (N, L, sigma), kinds, k, eta
    ("circle4", lambda: curve_mesh(1, 4), (4, 2, 0.15), (0, 1, 2, 3, 4, 5), 3.0, 3.0),
    ("kite5", lambda: curve_mesh(3, 5), (3, 2, 0.2), (2, 4), 7.5, 2.0),
    ("triangle6", lambda: polygon_mesh([(0, 0), (1.2, 0.1), (0.3, 0.9)], 2), (4, 2, 0.15), (1, 3),
     4.0, 4.0),
    ("arc4open", lambda: curve_mesh(5, 4, (1.3, 0.7, 0.3, 5.5), closed=False), (4, 2, 0.15), (4,),
     2.5, 1.5),
    ("circle3_q16", lambda: curve_mesh(1, 3), (16, 3, 0.15), (4,), 10.0, 10.0),
]


def make_entries():
    out = {}
    mp.mp.dps = 30
    for name, build, (N, L, sigma), kinds, k, eta in CASES:
        types, params, idx, nx, pv = build()
        out[name + "_type"] = types
        out[name + "_param"] = params
        out[name + "_index"] = idx
        out[name + "_next"] = nx
        out[name + "_prev"] = pv
        out[name + "_quad"] = np.array([N, L, sigma])
        out[name + "_keta"] = np.array([k, eta])
        for kind in kinds:
            if kind == 0:  # mass matrix (EvalIdent): diagonal sum of |Gamma'| w over Gauss1D(N)
                xt, wt = gauss_mp(N)
                A = np.zeros((len(types), len(types)), dtype=np.complex128)
                for e in range(len(types)):
                    acc = mp.mpf(0)
                    for t in range(N):
                        _, D = elem(types[e], params[e], xt[t])
                        acc += mp.sqrt(D[0] ** 2 + D[1] ** 2) * wt[t]
                    A[e, e] = float(acc)
            else:
                A = matrix_mp(types, params, idx, nx, pv, N, L, sigma, kind, k, eta)
            out["%s_kind%d" % (name, kind)] = A
            print(name, kind, file=sys.stderr, flush=True)
    np.savez_compressed(os.path.join(HERE, "entries_mpmath.npz"), **out)


CASES_COMPLEX = [
    # name, mesh builder, (N, L, sigma), kinds, k (complex), eta
    ("circle4_ck", lambda: curve_mesh(1, 4), (4, 2, 0.15), (1, 2, 3, 4), 6.0 - 1.5j, 6.0),
    ("triangle6_ck", lambda: polygon_mesh([(0, 0), (1.2, 0.1), (0.3, 0.9)], 2), (4, 2, 0.15),
     (1, 4), 4.0 + 2.0j, 4.0),
    ("kite5_ck", lambda: curve_mesh(3, 5), (3, 2, 0.2), (4,), 7.5 - 0.8j, 2.0),
]


def make_entries_complex():
    out = {}
    mp.mp.dps = 30
    for name, build, (N, L, sigma), kinds, k, eta in CASES_COMPLEX:
        types, params, idx, nx, pv = build()
        out[name + "_type"] = types
        out[name + "_param"] = params
        out[name + "_index"] = idx
        out[name + "_next"] = nx
        out[name + "_prev"] = pv
        out[name + "_quad"] = np.array([N, L, sigma])
        out[name + "_kc"] = np.array([k.real, k.imag, eta])
        for kind in kinds:
            out["%s_kind%d" % (name, kind)] = matrix_mp(types, params, idx, nx, pv, N, L, sigma,
                                                        kind, k, eta)
            print(name, kind, file=sys.stderr, flush=True)
    np.savez_compressed(os.path.join(HERE, "entries_complex_mpmath.npz"), **out)


def basis_mp(b, t):
    """phi_b(t) of a basis tuple (kind, degree, direction, k): 0 Const, 1 Pol, 2 Lin, 3 Wave,
    4 WavePol (lib/bem2d_basistypes.h:19-94)."""
    kind, deg, dr, k = b
    if kind == 0:
        return mp.mpc(1)
    if kind == 2:
        return mp.mpc(1 - t) if deg == -1 else mp.mpc(t)
    r = t ** deg if kind in (1, 4) else mp.mpf(1)
    if kind == 1:
        return mp.mpc(r)
    kk = mp.mpc(complex(k).real, complex(k).imag)
    return mp.exp(mp.mpc(0, 1) * kk * dr * t) * r


def matrix_mp_basis(types, params, index, nxt, prv, N, L, sigma, kind, k, eta, bases):
    """(n nb) x (n nb) Galerkin matrix with the bases (kind 0: the mass matrix)."""
    n, nb = len(types), len(bases)
    A = np.zeros((n * nb, n * nb), dtype=np.complex128)
    if kind == 0:
        xt, wt = gauss_mp(N)
        for e in range(n):
            acc = [[mp.mpc(0)] * nb for _ in range(nb)]
            for t in range(N):
                _, D = elem(types[e], params[e], xt[t])
                s = mp.sqrt(D[0] ** 2 + D[1] ** 2) * wt[t]
                for a in range(nb):
                    fa = mp.conj(basis_mp(bases[a], xt[t]))
                    for b in range(nb):
                        acc[a][b] += s * fa * basis_mp(bases[b], xt[t])
            for a in range(nb):
                for b in range(nb):
                    A[e * nb + a, e * nb + b] = complex(acc[a][b])
        return A
    k = mp.mpc(k.real, k.imag) if isinstance(k, complex) else mp.mpf(k)
    eta = mp.mpf(eta)
    rules = [rule_mp(w, N, L, sigma) for w in range(4)]
    for i in range(n):
        for j in range(n):
            acc = [[mp.mpc(0)] * nb for _ in range(nb)]
            if index[i] == index[j]:
                X, Y, W = rules[1]
                pts = []
                for q in range(len(X)):
                    for (s, t) in ((X[q], (1 - Y[q]) * X[q]), (1 - X[q], 1 - X[q] + Y[q] * X[q])):
                        pts.append((s, t, X[q] * W[q]))
            else:
                r = 0
                if nxt[i] == index[j]:
                    r = 3
                elif prv[i] == index[j]:
                    r = 2
                X, Y, W = rules[r]
                pts = [(X[q], Y[q], W[q]) for q in range(len(X))]
            for (s, t, w) in pts:
                Mx, Dx = elem(types[i], params[i], s)
                My, Dy = elem(types[j], params[j], t)
                n1, s1 = normal(Dx)
                n2, s2 = normal(Dy)
                g = kernel(kind, k, eta, Mx, My, n1, n2) * s1 * s2 * w
                fa = [mp.conj(basis_mp(bb, s)) for bb in bases]
                fb = [basis_mp(bb, t) for bb in bases]
                for a in range(nb):
                    for b in range(nb):
                        acc[a][b] += g * fa[a] * fb[b]
            for a in range(nb):
                for b in range(nb):
                    A[i * nb + a, j * nb + b] = complex(acc[a][b])
    return A


LIN = [(2, 1, 0.0, 0.0), (2, -1, 0.0, 0.0)]
POL2 = [(1, 0, 0.0, 0.0), (1, 1, 0.0, 0.0), (1, 2, 0.0, 0.0)]

CASES_BASIS = [
    # name, mesh builder, (N, L, sigma), kinds, k, eta, bases
    ("circle4_lin", lambda: curve_mesh(1, 4), (4, 2, 0.15), (0, 1, 3, 4), 3.0, 3.0, LIN),
    ("triangle6_pol2", lambda: polygon_mesh([(0, 0), (1.2, 0.1), (0.3, 0.9)], 2), (4, 2, 0.15),
     (0, 2), 4.0, 4.0, POL2),
    ("arc4open_wavepol1", lambda: curve_mesh(5, 4, (1.3, 0.7, 0.3, 5.5), closed=False),
     (4, 2, 0.15), (0, 4), 2.5, 1.5,
     [(4, 0, 1.0, 1.7), (4, 0, -1.0, 1.7), (4, 1, 1.0, 1.7), (4, 1, -1.0, 1.7)]),
    ("circle3_wave_ck", lambda: curve_mesh(1, 3), (4, 2, 0.15), (1, 4), 6.0 - 1.5j, 6.0,
     [(3, 0, 1.0, 2.0 - 0.5j), (3, 0, -1.0, 2.0 - 0.5j)]),
]


def make_entries_basis():
    out = {}
    mp.mp.dps = 30
    for name, build, (N, L, sigma), kinds, k, eta, bases in CASES_BASIS:
        types, params, idx, nx, pv = build()
        out[name + "_type"] = types
        out[name + "_param"] = params
        out[name + "_index"] = idx
        out[name + "_next"] = nx
        out[name + "_prev"] = pv
        out[name + "_quad"] = np.array([N, L, sigma])
        out[name + "_kc"] = np.array([complex(k).real, complex(k).imag, eta])
        out[name + "_bases"] = np.array([[b[0], b[1], b[2], complex(b[3]).real,
                                          complex(b[3]).imag] for b in bases])
        for kind in kinds:
            out["%s_kind%d" % (name, kind)] = matrix_mp_basis(types, params, idx, nx, pv, N, L,
                                                              sigma, kind, k, eta, bases)
            print(name, kind, file=sys.stderr, flush=True)
    np.savez_compressed(os.path.join(HERE, "entries_basis_mpmath.npz"), **out)


def make_hankel_complex():
    zs = []
    for ph in np.linspace(-0.95 * np.pi / 2, 0.95 * np.pi / 2, 39):
        r = list(np.logspace(-6, np.log10(2e3), 60))
        for c in (2.0, 20.0):
            r += list(c * (1 + np.linspace(-1e-3, 1e-3, 5)))
        zs += [x * np.exp(1j * ph) for x in r]
    # H^(1) is recessive above the real axis (~e^-Im z); the 30-digit values there need
    # 2 Im z / ln 10 extra digits, so the fixture stops at Im z = 120 (|H| ~ 1e-53)
    zs = np.array([z for z in zs if -650 < z.imag < 120])
    vals = np.zeros((zs.size, 4))
    for i, z in enumerate(zs):
        Z = mp.mpc(z.real, z.imag)
        with mp.workdps(30 + int(np.ceil(2 * max(z.imag, 0) / np.log(10)))):
            h0, h1 = mp.hankel1(0, Z), mp.hankel1(1, Z)
        vals[i] = [float(h0.real), float(h0.imag), float(h1.real), float(h1.imag)]
    np.savez_compressed(os.path.join(HERE, "hankel_complex_mpmath.npz"), z=zs, h0h1=vals)


def make_hankel():
    mp.mp.dps = 50
    xs = list(np.logspace(-8, np.log10(5e4), 2500))
    for c in (2.0, 25.0, 8.0, 12.0, 16.0, 20.0, 3.0, 4.0, 5.0, 6.0, 10.0):
        xs += list(c * (1 + np.linspace(-1e-3, 1e-3, 21)))
    for f in (mp.besselj, mp.bessely):
        for nu in (0, 1):
            for m in range(1, 8):
                z = float(mp.besseljzero(nu, m) if f is mp.besselj else mp.besselyzero(nu, m))
                xs += list(z * (1 + np.linspace(-1e-6, 1e-6, 5)))
    xs = np.unique(np.array(xs, dtype=np.float64))
    vals = np.zeros((xs.size, 4))
    for i, x in enumerate(xs):
        X = mp.mpf(float(x))
        vals[i] = [float(mp.besselj(0, X)), float(mp.bessely(0, X)), float(mp.besselj(1, X)),
                   float(mp.bessely(1, X))]
    np.savez_compressed(os.path.join(HERE, "hankel_mpmath.npz"), x=xs, j0y0j1y1=vals)


def make_gauss():
    out = {}
    for n in (1, 2, 3, 4, 5, 8, 16, 30):
        x, w = gauss_mp(n)
        out["x%d" % n] = np.array([float(v) for v in x])
        out["w%d" % n] = np.array([float(v) for v in w])
    np.savez_compressed(os.path.join(HERE, "gauss_mpmath.npz"), **out)


if __name__ == "__main__":
    what = sys.argv[1:] or ["gauss", "hankel", "entries", "hankel_complex", "entries_complex",
                            "entries_basis"]
    if "gauss" in what:
        make_gauss()
    if "hankel" in what:
        make_hankel()
    if "entries" in what:
        make_entries()
    if "hankel_complex" in what:
        make_hankel_complex()
    if "entries_complex" in what:
        make_entries_complex()
    if "entries_basis" in what:
        make_entries_basis()
