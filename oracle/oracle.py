"""ctypes front end of the CPU oracle (oracle/liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke entry and bench.py's cpu_baseline leg import this module; the product
package never does.  Every function here forwards to the C restatement in bem2d_oracle.c, which
cites the reference file:line it follows.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

R_GAUSS2D, R_AG1, R_AG2, R_AG3 = 0, 1, 2, 3
K_NONE, K_SL, K_DL, K_CDL, K_COMBINED, K_ID = 0, 1, 2, 3, 4, 5
B_CONST, B_POL, B_LIN, B_WAVE, B_WAVEPOL = 0, 1, 2, 3, 4
F_PLANE, F_NORMAL_PLANE, F_COMBINED_PLANE = 0, 1, 2


def lib(variant="plain"):
    """The oracle library: "plain" (IEEE multiply/add and the C library's sin / cos, as the
    reference's x86-64 build), "fma" (the same source with fused multiply-adds) or "crtrig" (plain
    arithmetic with correctly rounded curve sin / cos); the last two measure the noise floor."""
    if variant not in _LIBS:
        name = {"plain": "liboracle.so", "fma": "liboracle_fma.so",
                "crtrig": "liboracle_crtrig.so"}[variant]
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s is not built (run make -C oracle)" % name)
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        d = ctypes.c_double
        i32 = ctypes.c_int
        L.oracle_gauss.argtypes = [i32, P, P]
        L.oracle_rule_size.argtypes = [i32, i32, i32]
        L.oracle_rule.argtypes = [i32, i32, i32, d, P, P, P]
        L.oracle_hankel_many.argtypes = [P, i64, P]
        L.oracle_hankel_many.restype = None
        L.oracle_assemble.argtypes = [i64, P, P, P, P, P, i32, i32, d, i32, d, d, d, d, P, i64,
                                      P, i64]
        L.oracle_rhs.argtypes = [i64, P, P, i32, i32, d, d, P, i32, d, P, i64]
        L.oracle_lu_solve.argtypes = [i64, i64, P, P]
        L.oracle_eval_sl.argtypes = [i64, P, P, i32, d, P, P, i64, P]
        L.oracle_farfield.argtypes = [i64, P, P, i32, d, P, i32, P, i64, P]
        L.oracle_assemble_ck.argtypes = [i64, P, P, P, P, P, i32, i32, d, i32, d, d, d, d, d,
                                         P, i64, P, i64]
        L.oracle_rhs_ck.argtypes = [i64, P, P, i32, i32, d, d, d, P, i32, d, P, i64]
        L.oracle_eval_sl_ck.argtypes = [i64, P, P, i32, d, d, P, P, i64, P]
        L.oracle_hankel_complex_many.argtypes = [P, i64, P]
        L.oracle_hankel_complex_many.restype = None
        L.oracle_elem_eval.argtypes = [i64, P, P, i64, d, P]
        L.oracle_elem_eval.restype = None
        L.oracle_assemble_basis.argtypes = [i64, P, P, P, P, P, i32, P, P, P, P, i32, i32, d,
                                            i32, d, d, d, d, d, P, i64, P, i64]
        L.oracle_rhs_basis.argtypes = [i64, P, P, i32, P, P, P, P, i32, i32, d, d, d, P, i32, d,
                                       P, i64]
        L.oracle_eval_sl_basis.argtypes = [i64, P, P, i32, P, P, P, P, i32, d, d, P, P, i64, P]
        L.oracle_farfield_basis.argtypes = [i64, P, P, i32, P, P, P, P, i32, d, P, i32, P, i64,
                                            P]
        L.oracle_curve_length.argtypes = [i32, P]
        L.oracle_curve_length.restype = d
        L.oracle_arclength_partition.argtypes = [i32, P, i32, P]
        _LIBS[variant] = L
    return _LIBS[variant]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    rc = lib().oracle_gauss(n, _p(x), _p(w))
    if rc:
        raise ValueError("oracle_gauss failed")
    return x, w


def rule(which, N, L=3, sigma=0.15):
    m = lib().oracle_rule_size(which, N, L)
    x, y, w = np.zeros(m), np.zeros(m), np.zeros(m)
    c = lib().oracle_rule(which, N, L, sigma, _p(x), _p(y), _p(w))
    assert c == m
    return x, y, w


def hankel(x):
    """Returns (J0, Y0, J1, Y1) arrays of real x > 0."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((x.size, 4))
    lib().oracle_hankel_many(_p(x), x.size, _p(out))
    return out[:, 0], out[:, 1], out[:, 2], out[:, 3]


def hankel_complex(z):
    """Returns (H0, H1) arrays (complex128) of H0^(1), H1^(1) at complex z."""
    z = np.ascontiguousarray(z, dtype=np.complex128)
    zz = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1).reshape(-1))
    out = np.zeros((z.size, 4))
    lib().oracle_hankel_complex_many(_p(zz), z.size, _p(out))
    return out[:, 0] + 1j * out[:, 1], out[:, 2] + 1j * out[:, 3]


def _kparts(k):
    """(Re k, Im k) of a real or complex wavenumber (freqtype {re, im})."""
    k = complex(k)
    return k.real, k.imag


def assemble(geom, N, L, sigma, kind, k, eta, mass_scale=0.0, kernel_scale=1.0, cols=None,
             variant="plain"):
    """Column-major N x ncols complex128 block of mass_scale*M + kernel_scale*K; k may be complex."""
    n = geom.n
    if cols is None:
        ncols = n
        cptr = None
    else:
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        ncols = cols.size
        cptr = _p(cols)
    out = np.zeros((ncols, n), dtype=np.complex128)  # row c = column cols[c]
    kr, ki = _kparts(k)
    rc = lib(variant).oracle_assemble_ck(n, _p(geom.type), _p(geom.param), _p(geom.index),
                                         _p(geom.next), _p(geom.prev), N, L, sigma, kind, kr, ki,
                                         eta, mass_scale, kernel_scale, cptr, ncols, _p(out), n)
    if rc:
        raise ValueError("oracle_assemble: bad quadrature option")
    return out.T  # (n, ncols) view, column-major storage


def rhs(geom, N, kind, k, eta, dirs, scale=1.0):
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 2)
    nrhs = dirs.shape[0]
    out = np.zeros((nrhs, geom.n), dtype=np.complex128)
    kr, ki = _kparts(k)
    lib().oracle_rhs_ck(geom.n, _p(geom.type), _p(geom.param), N, kind, kr, ki, eta, _p(dirs),
                        nrhs, scale, _p(out), geom.n)
    return out.T


def lu_solve(A, B):
    """X = A^{-1} B with zgetrf/zgetrs semantics; raises on a singular matrix (LapackError)."""
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.complex128)
    Bf = np.array(B, dtype=np.complex128, order="F", copy=True)
    if Bf.ndim == 1:
        Bf = Bf.reshape(n, 1, order="F")
    info = lib().oracle_lu_solve(n, Bf.shape[1], _p(Af), _p(Bf))
    if info:
        raise np.linalg.LinAlgError("singular matrix: U(%d,%d) = 0" % (info, info))
    return Bf


def eval_sl(geom, N, k, coef, pts):
    coef = np.ascontiguousarray(coef, dtype=np.complex128)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(pts.shape[0], dtype=np.complex128)
    kr, ki = _kparts(k)
    lib().oracle_eval_sl_ck(geom.n, _p(geom.type), _p(geom.param), N, kr, ki, _p(coef), _p(pts),
                            pts.shape[0], _p(out))
    return out


def farfield(geom, N, k, coef, theta):
    coef = np.asfortranarray(coef, dtype=np.complex128)
    if coef.ndim == 1:
        coef = coef.reshape(-1, 1, order="F")
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    nrhs = coef.shape[1]
    out = np.zeros((nrhs, theta.size), dtype=np.complex128)
    lib().oracle_farfield(geom.n, _p(geom.type), _p(geom.param), N, k, _p(coef), nrhs, _p(theta),
                          theta.size, _p(out))
    return out.T


def elem_eval(geom, e, t):
    out = np.zeros(6)
    lib().oracle_elem_eval(geom.n, _p(geom.type), _p(geom.param), e, t, _p(out))
    return out


def noise_floor(geom, N, L, sigma, kind, k, eta, mass_scale=0.0, kernel_scale=1.0, cols=None):
    """How far rounding alone moves the reference's own matrix for this input: the larger relative
    Frobenius distance of the plain oracle from the oracle with fused multiply-adds and from the
    oracle with correctly rounded curve sin / cos (the reference's matrix depends on the libm it is
    linked against: glibc's sin and cos are one ulp off in about 0.1% of arguments, and near-field
    double-layer entries amplify a node position's last bit by ~1 / (curvature d^2))."""
    a = assemble(geom, N, L, sigma, kind, k, eta, mass_scale, kernel_scale, cols)
    na = np.linalg.norm(a)
    if na == 0:
        return 0.0
    out = 0.0
    for v in ("fma", "crtrig"):
        b = assemble(geom, N, L, sigma, kind, k, eta, mass_scale, kernel_scale, cols, variant=v)
        out = max(out, float(np.linalg.norm(a - b) / na))
    return out


def _bases(bases):
    """Arrays of a basis list [(kind, deg, dir, k), ...] (kind B_*; deg: the power of PolBasis and
    WavePolBasis, the sign +1 / -1 of LinBasis; dir and k of WaveBasis / WavePolBasis)."""
    kind = np.ascontiguousarray([b[0] for b in bases], dtype=np.int32)
    deg = np.ascontiguousarray([b[1] for b in bases], dtype=np.int32)
    dr = np.ascontiguousarray([b[2] for b in bases], dtype=np.float64)
    kk = np.ascontiguousarray([[complex(b[3]).real, complex(b[3]).imag] for b in bases],
                              dtype=np.float64).reshape(-1)
    return kind, deg, dr, kk


def assemble_basis(geom, bases, N, L, sigma, kind, k, eta, mass_scale=0.0, kernel_scale=1.0,
                   cols=None, variant="plain"):
    """(n nb) x ncols block with nb basis functions per element, row = element * nb + basis."""
    nb = len(bases)
    nn = geom.n * nb
    if cols is None:
        ncols, cptr = nn, None
    else:
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        ncols, cptr = cols.size, _p(cols)
    bk, bd, bdir, bkk = _bases(bases)
    out = np.zeros((ncols, nn), dtype=np.complex128)
    kr, ki = _kparts(k)
    rc = lib(variant).oracle_assemble_basis(geom.n, _p(geom.type), _p(geom.param), _p(geom.index),
                                     _p(geom.next), _p(geom.prev), nb, _p(bk), _p(bd), _p(bdir),
                                     _p(bkk), N, L, sigma, kind, kr, ki, eta, mass_scale,
                                     kernel_scale, cptr, ncols, _p(out), nn)
    if rc:
        raise ValueError("oracle_assemble_basis: bad option")
    return out.T


def rhs_basis(geom, bases, N, kind, k, eta, dirs, scale=1.0):
    nb = len(bases)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 2)
    nrhs = dirs.shape[0]
    out = np.zeros((nrhs, geom.n * nb), dtype=np.complex128)
    bk, bd, bdir, bkk = _bases(bases)
    kr, ki = _kparts(k)
    lib().oracle_rhs_basis(geom.n, _p(geom.type), _p(geom.param), nb, _p(bk), _p(bd), _p(bdir),
                           _p(bkk), N, kind, kr, ki, eta, _p(dirs), nrhs, scale, _p(out),
                           geom.n * nb)
    return out.T


def eval_sl_basis(geom, bases, N, k, coef, pts):
    nb = len(bases)
    coef = np.ascontiguousarray(coef, dtype=np.complex128)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(pts.shape[0], dtype=np.complex128)
    bk, bd, bdir, bkk = _bases(bases)
    kr, ki = _kparts(k)
    lib().oracle_eval_sl_basis(geom.n, _p(geom.type), _p(geom.param), nb, _p(bk), _p(bd),
                               _p(bdir), _p(bkk), N, kr, ki, _p(coef), _p(pts), pts.shape[0],
                               _p(out))
    return out


def farfield_basis(geom, bases, N, k, coef, theta):
    nb = len(bases)
    coef = np.asfortranarray(coef, dtype=np.complex128)
    if coef.ndim == 1:
        coef = coef.reshape(-1, 1, order="F")
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    nrhs = coef.shape[1]
    out = np.zeros((nrhs, theta.size), dtype=np.complex128)
    bk, bd, bdir, bkk = _bases(bases)
    lib().oracle_farfield_basis(geom.n, _p(geom.type), _p(geom.param), nb, _p(bk), _p(bd),
                                _p(bdir), _p(bkk), N, float(k), _p(coef), nrhs, _p(theta),
                                theta.size, _p(out))
    return out.T


def noise_floor_basis(geom, bases, N, L, sigma, kind, k, eta, mass_scale=0.0, kernel_scale=1.0,
                      cols=None):
    """noise_floor for the matrix with bases."""
    a = assemble_basis(geom, bases, N, L, sigma, kind, k, eta, mass_scale, kernel_scale, cols)
    na = np.linalg.norm(a)
    if na == 0:
        return 0.0
    out = 0.0
    for v in ("fma", "crtrig"):
        b = assemble_basis(geom, bases, N, L, sigma, kind, k, eta, mass_scale, kernel_scale, cols,
                           variant=v)
        out = max(out, float(np.linalg.norm(a - b) / na))
    return out


def curve_length(ctype, c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    return lib().oracle_curve_length(ctype, _p(c))


def arclength_partition(ctype, c, n):
    """Parameters t_0 = 0 < ... < t_n = 1 splitting the curve into n pieces of equal arc length."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    t = np.zeros(n + 1)
    if lib().oracle_arclength_partition(ctype, _p(c), n, _p(t)):
        raise ValueError("arclength_partition: n < 1")
    return t
