"""Generates hankel_tables.h: the coefficient tables of the device H0^(1)/H1^(1) of real argument.

    python bem2d_amd/csrc/gen_hankel_tables.py      (needs mpmath; the header is committed)

Representation (x > 0):
  x < 2      ascending series in t = x^2/4 (A&S 9.1.10, 9.1.11 with n = 0, 1):
               J0 = S_J0(t),  J1 = (x/2) S_J1(t),
               Y0 = (2/pi) ln(x/2) J0 + S_Y0(t),
               Y1 = (2/pi) ln(x/2) J1 - 2/(pi x) + (x/2) S_Y1(t)
  x >= 2     modulus/phase form of A&S 9.2.5-9.2.6:
               H_nu(x) = sqrt(2/(pi x)) (P_nu(x) + i Q_nu(x)) exp(i (x - (2 nu + 1) pi/4)),
             with P_nu and x*Q_nu fitted per piece [a, b) as polynomials of degree DEGS[p] in
             z = (y - ym)/yh in [-1, 1], y = (a/x)^2 (the last piece reaches x = infinity, y -> 0).
The header gives every table as an initializer macro, so that the device code can hold them in
__constant__ memory (read through scalar loads into SGPR operands) and the host code in constexpr
arrays.  Two more tables serve the far-field fast path of the assembly:
  FAR   [FAR_X, inf) = [40, inf), where most far-field pairs of an assembly fall, as plain
        polynomials of degree 4 in w = 1/x^2 (the affine map to z folded into the coefficients);
  TRIG  sin r = r + r^3 S(r^2), cos r = 1 + r^2 C(r^2) on |r| <= pi/4, S of degree 5 and C of
        degree 6, interpolated at Chebyshev nodes in r^2 (near-minimax; 1.4e-17 relative and
        1.2e-19 absolute), plus 2/pi and pi/2 split into three doubles for the reduction;
  TRIG2 the same on |r| <= pi/2 (the half-turn reduction of the Hankel phase), S and C of degree 7
        (1.3e-18 and 1.6e-17 absolute), plus 1/pi and pi split into three doubles.
The fits are Chebyshev interpolants computed at 40 digits from mpmath's hankel1 and converted
to the monomial basis in z; the script reports the float64 Horner error of every table.
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))
PIECES = [2.0, 2.6, 3.5, 5.0, 8.0]  # piece p covers [PIECES[p], PIECES[p+1]); last [8, inf)
DEGS = [12, 12, 12, 12, 10]         # polynomial degree per piece
FAR_X, FAR_DEG = 40.0, 4            # the far table: [FAR_X, inf), polynomials in w = 1/x^2
DEG = max(DEGS)
NSER = 14  # series terms k = 0..NSER-1


def PQ(nu, x):
    x = mp.mpf(x)
    h = mp.hankel1(nu, x)
    v = h * mp.exp(-1j * (x - (2 * nu + 1) * mp.pi / 4)) * mp.sqrt(mp.pi * x / 2)
    return v.real, v.imag


def cheb_to_monomial(c):
    """Chebyshev coefficients c_k (sum c_k T_k(z)) -> monomial coefficients in z (ascending)."""
    n = len(c)
    T = [[mp.mpf(1)] + [mp.mpf(0)] * (n - 1), [mp.mpf(0), mp.mpf(1)] + [mp.mpf(0)] * (n - 2)]
    for k in range(2, n):
        t = [mp.mpf(0)] * n
        for i in range(n - 1):
            t[i + 1] += 2 * T[k - 1][i]
        for i in range(n):
            t[i] -= T[k - 2][i]
        T.append(t)
    m = [mp.mpf(0)] * n
    for k in range(n):
        for i in range(n):
            m[i] += c[k] * T[k][i]
    return m


def fit_piece(f, ylo, yhi, deg):
    ym, yh = (ylo + yhi) / 2, (yhi - ylo) / 2
    nodes = [mp.cos(mp.pi * (j + mp.mpf("0.5")) / (deg + 1)) for j in range(deg + 1)]
    vals = [f(ym + yh * z) for z in nodes]
    c = []
    for k in range(deg + 1):
        s = mp.fsum(vals[j] * mp.cos(mp.pi * k * (j + mp.mpf("0.5")) / (deg + 1))
                    for j in range(deg + 1))
        c.append(2 * s / (deg + 1))
    c[0] /= 2
    return cheb_to_monomial(c), ym, yh


def horner(coefs, z):
    r = np.float64(coefs[-1])
    for cc in coefs[-2::-1]:
        r = r * z + np.float64(cc)
    return r


def fmt(v):
    return repr(float(v))


def main():
    lines = []
    emit = lines.append
    emit("// Generated by gen_hankel_tables.py -- do not edit.  Coefficients of the device")
    emit("// H0^(1), H1^(1) of real argument; see the generator for the representation.")
    emit("#pragma once")
    emit("")
    emit("#define BEM2D_HANKEL_NPIECE %d" % len(PIECES))
    emit("#define BEM2D_HANKEL_DEGMAX %d" % DEG)
    emit("#define BEM2D_HANKEL_NSER %d" % NSER)
    emit("")
    # ---- series
    g = mp.euler
    H = [mp.mpf(0)]
    for k in range(1, NSER + 2):
        H.append(H[-1] + mp.mpf(1) / k)
    sj0 = [(-1) ** k / mp.factorial(k) ** 2 for k in range(NSER)]
    sj1 = [(-1) ** k / (mp.factorial(k) * mp.factorial(k + 1)) for k in range(NSER)]
    sy0 = [(2 / mp.pi) * (-1) ** k * (g - H[k]) / mp.factorial(k) ** 2 for k in range(NSER)]
    sy1 = [-(1 / mp.pi) * (-1) ** k * (H[k] + H[k + 1] - 2 * g) /
           (mp.factorial(k) * mp.factorial(k + 1)) for k in range(NSER)]
    emit("// series tables [J0, J1, Y0, Y1][k], ascending powers of t = x^2/4")
    emit("#define BEM2D_HANKEL_SER_INIT { \\")
    for arr in (sj0, sj1, sy0, sy1):
        emit("    {" + ", ".join(fmt(v) for v in arr) + "}, \\")
    emit("}")
    # ---- pieces
    fits = []
    zmap = []
    worst = 0.0
    for p, a in enumerate(PIECES):
        a = mp.mpf(a)
        b = mp.mpf(PIECES[p + 1]) if p + 1 < len(PIECES) else None
        ylo = (a / b) ** 2 if b is not None else mp.mpf(0)
        deg = DEGS[p]
        row = []
        for nu in (0, 1):
            def fP(y, nu=nu):
                if y == 0:
                    return mp.mpf(1)
                return PQ(nu, a / mp.sqrt(y))[0]

            def fQ(y, nu=nu):
                if y == 0:
                    return (4 * nu * nu - 1) / mp.mpf(8)
                x = a / mp.sqrt(y)
                return PQ(nu, x)[1] * x
            for fname, f in (("P%d" % nu, fP), ("Q%d" % nu, fQ)):
                coefs, ym, yh = fit_piece(f, ylo, mp.mpf(1), deg)
                cf = [float(c) for c in coefs]
                row.append(cf + [0.0] * (DEG - deg))
                err = 0.0
                ys = np.linspace(float(ylo), 1.0, 257)
                if b is None:
                    ys = ys[1:]
                for y in ys:
                    z = (np.float64(y) - np.float64(ym)) / np.float64(yh)
                    approx = horner(cf, z)
                    exact = f(mp.mpf(y))
                    scale = 1 if fname[0] == "P" else a / mp.sqrt(mp.mpf(y))  # Q = fit / x
                    err = max(err, float(abs((approx - exact) / scale)))
                worst = max(worst, err)
                print("piece %d [%s, %s) deg %d %s: max |err| = %.3g" % (p, a, b, deg, fname, err))
        fits.append(row)
        zmap.append((float(a) ** 2, float(ym), 1.0 / float(yh)))
    emit("// piece p covers x in [PIECE_X[p], PIECE_X[p+1]); polynomial degree PIECE_DEG[p]")
    emit("static constexpr double BEM2D_PIECE_X[BEM2D_HANKEL_NPIECE] = {%s};" %
         ", ".join(fmt(v) for v in PIECES))
    emit("static constexpr int BEM2D_PIECE_DEG[BEM2D_HANKEL_NPIECE] = {%s};" %
         ", ".join(str(d) for d in DEGS))
    emit("// [piece][a^2, ym, 1/yh]: y = a^2/x^2, z = (y - ym)/yh")
    emit("#define BEM2D_HANKEL_ZMAP_INIT { \\")
    for z in zmap:
        emit("    {%s}, \\" % ", ".join(fmt(v) for v in z))
    emit("}")
    emit("// [piece][P0, xQ0, P1, xQ1][ascending powers of z]")
    emit("#define BEM2D_HANKEL_FIT_INIT { \\")
    for row in fits:
        emit("    { \\")
        for cf in row:
            emit("        {%s}, \\" % ", ".join(fmt(v) for v in cf))
        emit("    }, \\")
    emit("}")
    # ---- far piece as polynomials in w = 1/x^2: z = (a^2 w - ym)/yh ------------------------------
    a = mp.mpf(FAR_X)
    degf = FAR_DEG
    far = []
    for nu in (0, 1):
        def fP(y, nu=nu):
            return mp.mpf(1) if y == 0 else PQ(nu, a / mp.sqrt(y))[0]

        def fQ(y, nu=nu):
            if y == 0:
                return (4 * nu * nu - 1) / mp.mpf(8)
            x = a / mp.sqrt(y)
            return PQ(nu, x)[1] * x
        for f in (fP, fQ):
            cz, ym, yh = fit_piece(f, mp.mpf(0), mp.mpf(1), degf)
            # p(z) with z = (a^2 w - ym)/yh  ->  coefficients in w
            cw = [mp.mpf(0)] * (degf + 1)
            lin = [-ym / yh, a * a / yh]  # z = lin[0] + lin[1] w
            powz = [mp.mpf(1)] + [mp.mpf(0)] * degf
            for kk in range(degf + 1):
                for i in range(degf + 1):
                    cw[i] += cz[kk] * powz[i]
                nxt = [mp.mpf(0)] * (degf + 1)
                for i in range(degf):
                    nxt[i] += powz[i] * lin[0]
                    nxt[i + 1] += powz[i] * lin[1]
                powz = nxt
            far.append([float(c) for c in cw])
            err = 0.0
            for x in np.linspace(float(a), 4000.0, 400):
                approx = horner([float(c) for c in cw], 1.0 / (x * x))
                exact = f(a * a / mp.mpf(x) ** 2)
                scale = 1 if f is fP else x
                err = max(err, float(abs((approx - exact) / scale)))
            worst = max(worst, err)
            print("far piece w-form: max |err| = %.3g" % err)
    emit("// far table [%s, inf): [P0, xQ0, P1, xQ1][ascending powers of w = 1/x^2]" % FAR_X)
    emit("#define BEM2D_HANKEL_FAR_X %s" % fmt(FAR_X))
    emit("#define BEM2D_HANKEL_FAR_DEG %d" % degf)
    emit("#define BEM2D_HANKEL_FAR_INIT { \\")
    for cf in far:
        emit("    {%s}, \\" % ", ".join(fmt(v) for v in cf))
    emit("}")
    # ---- trig kernels -----------------------------------------------------------------------------
    R2 = (mp.pi / 4) ** 2

    def interp(fn, deg, R2=R2):
        nodes = [R2 / 2 + R2 / 2 * mp.cos(mp.pi * (j + mp.mpf("0.5")) / (deg + 1))
                 for j in range(deg + 1)]
        A = mp.matrix([[u ** kk for kk in range(deg + 1)] for u in nodes])
        yv = mp.matrix([fn(u) for u in nodes])
        c = mp.lu_solve(A, yv)
        return [c[kk] for kk in range(deg + 1)]
    Sc = interp(lambda u: (mp.sin(mp.sqrt(u)) - mp.sqrt(u)) / (u * mp.sqrt(u)), 5)
    Cc = interp(lambda u: (mp.cos(mp.sqrt(u)) - 1) / u, 6)
    with mp.workdps(80):  # P3 is 2^-110 below pi/2: the split needs more than 40 digits
        pio2 = mp.pi / 2
        P1 = float(pio2)
        P2 = float(pio2 - mp.mpf(P1))
        P3 = float(pio2 - mp.mpf(P1) - mp.mpf(P2))
        twoopi = float(2 / mp.pi)
    trig = [twoopi, P1, P2, P3] + [float(c) for c in reversed(Sc)] + \
        [float(c) for c in reversed(Cc)]
    emit("// [2/pi, P1, P2, P3, S5..S0, C6..C0]: sin r = r + r^3 S(r^2), cos r = 1 + r^2 C(r^2)")
    emit("#define BEM2D_TRIG_NS 6")
    emit("#define BEM2D_TRIG_NC 7")
    emit("#define BEM2D_TRIG_INIT {%s}" % ", ".join(fmt(v) for v in trig))
    # half-turn kernels: |r| <= pi/2 (a hair beyond, for the rounding of the reduction)
    H2 = (mp.pi / 2 * (1 + mp.mpf("1e-9"))) ** 2
    S2 = interp(lambda u: (mp.sin(mp.sqrt(u)) - mp.sqrt(u)) / (u * mp.sqrt(u)), 7, H2)
    C2 = interp(lambda u: (mp.cos(mp.sqrt(u)) - 1) / u, 7, H2)
    for name, fn, cf, scale in (("sin", lambda u: (mp.sin(mp.sqrt(u)) - mp.sqrt(u)) / (u * mp.sqrt(u)),
                                 S2, lambda u: u * mp.sqrt(u)),
                                ("cos", lambda u: (mp.cos(mp.sqrt(u)) - 1) / u, C2, lambda u: u)):
        err = max(abs((fn(u) - sum(cf[kk] * u ** kk for kk in range(len(cf)))) * scale(u))
                  for u in [H2 * mp.mpf(t) for t in np.linspace(1e-3, 1.0, 400)])
        print("half-turn %s: max |err| = %.3g" % (name, float(err)))
    with mp.workdps(80):
        PI1 = float(mp.pi)
        PI2 = float(mp.pi - mp.mpf(PI1))
        PI3 = float(mp.pi - mp.mpf(PI1) - mp.mpf(PI2))
        oopi = float(1 / mp.pi)
    trig2 = [oopi, PI1, PI2, PI3] + [float(c) for c in reversed(S2)] + \
        [float(c) for c in reversed(C2)]
    emit("// [1/pi, PI1, PI2, PI3, S7..S0, C7..C0]: the same on |r| <= pi/2")
    emit("#define BEM2D_TRIG2_NS 8")
    emit("#define BEM2D_TRIG2_NC 8")
    emit("#define BEM2D_TRIG2_INIT {%s}" % ", ".join(fmt(v) for v in trig2))
    emit("// worst float64 Horner error of the fitted tables (absolute, on P and Q): %.3g" % worst)
    with open(os.path.join(HERE, "hankel_tables.h"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("worst", worst)


if __name__ == "__main__":
    main()
