"""ctypes binding of the C-ABI in include/bem2d_hip.h (libbem2d_hip.so, built in-tree).

This is plumbing for the tests, the smoke entry and bench.py: device buffers are torch tensors
(complex128, column-major views), passed to the library as raw pointers.  Loading fails loudly
when the HIP library has not been built -- there is no fallback.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbem2d_hip.so")

OK, ERR_ARG, ERR_SINGULAR, ERR_NOMEM, ERR_DEVICE, ERR_UNSUPPORTED, ERR_NOT_CONVERGED = range(7)
ELEM_FLAT, ELEM_CIRCLE, ELEM_TREFOIL, ELEM_KITE, ELEM_ELLIPSE, ELEM_ELLIPSE_ARC, ELEM_INV_ELLIPSE = \
    range(7)
KERNEL_NONE, KERNEL_SL, KERNEL_DL, KERNEL_CDL, KERNEL_COMBINED, KERNEL_ID = range(6)
FIELD_PLANE, FIELD_NORMAL_PLANE, FIELD_COMBINED_PLANE = range(3)
PROF_ASSEMBLE_REGULAR, PROF_ASSEMBLE_SPECIAL, PROF_LU, PROF_LU_UPDATE = range(4)
BASIS_CONST, BASIS_POL, BASIS_LIN, BASIS_WAVE, BASIS_WAVEPOL = range(5)
MAX_BASES = 8

# Exported symbols, with (restype, argtypes).  Kept in the order of include/bem2d_hip.h.
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_I = ctypes.c_int


class Geometry(ctypes.Structure):
    _fields_ = [("n", _I64), ("type", _P), ("param", _P), ("index", _P), ("next", _P),
                ("prev", _P)]


class QuadOption(ctypes.Structure):
    _fields_ = [("L", _I), ("N", _I), ("sigma", _D)]


class Basis(ctypes.Structure):
    """bem2d_basis: one basis function per element (kind BASIS_*; degree = the power of POL /
    WAVEPOL, the sign of LIN; direction and k of WAVE / WAVEPOL)."""
    _fields_ = [("kind", _I), ("degree", _I), ("direction", _D), ("k_re", _D), ("k_im", _D)]


def const_basis():
    """ConstBasis (lib/bem2d_basistypes.h:16-25)."""
    return [(BASIS_CONST, 0, 0.0, 0.0)]


def pol_bases(degree):
    """PolBasis::AddBasis(degree, geom): t^0 .. t^degree (lib/bem2d_basistypes.cpp:12-17)."""
    return [(BASIS_POL, d, 0.0, 0.0) for d in range(degree + 1)]


def lin_bases():
    """LinBasis::AddBasis(geom): t, then 1 - t (lib/bem2d_basistypes.cpp:21-26)."""
    return [(BASIS_LIN, 1, 0.0, 0.0), (BASIS_LIN, -1, 0.0, 0.0)]


def wave_bases(k):
    """WaveBasis::AddBasis(k, geom): exp(ikt), exp(-ikt) (lib/bem2d_basistypes.cpp:30-34)."""
    return [(BASIS_WAVE, 0, 1.0, k), (BASIS_WAVE, 0, -1.0, k)]


def wavepol_bases(degree, k):
    """WavePolBasis::AddBasis(degree, k, geom): exp(+-ikt) t^d, d = 0 .. degree, the + direction
    first for each d (lib/bem2d_basistypes.cpp:38-44)."""
    out = []
    for d in range(degree + 1):
        out += [(BASIS_WAVEPOL, d, 1.0, k), (BASIS_WAVEPOL, d, -1.0, k)]
    return out


def basis_array(bases):
    """ctypes array of bem2d_basis from a list of (kind, degree, direction, k) tuples."""
    arr = (Basis * len(bases))()
    for i, (kind, deg, dr, k) in enumerate(bases):
        kc = complex(k)
        arr[i] = Basis(int(kind), int(deg), float(dr), kc.real, kc.imag)
    return arr


def basis_values(bases, t):
    """phi_b(t_q) on the host: (len(t), nb) complex array."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros((t.size, len(bases), 2))
    check(lib().bem2d_basis_values(basis_array(bases), len(bases), _np_ptr(t), t.size,
                                   _np_ptr(out)))
    return out[..., 0] + 1j * out[..., 1]


def curve_length(curve_type, args, t0=0.0, t1=1.0):
    """Curve::Length(t0, t1) of curve type ELEM_CIRCLE .. ELEM_INV_ELLIPSE."""
    a = np.zeros(4)
    a[:len(args)] = args
    out = _D()
    check(lib().bem2d_curve_length(int(curve_type), _np_ptr(a), t0, t1, ctypes.byref(out)))
    return out.value


def arclength_params(curve_type, args, s):
    """Curve parameters t_i at arc lengths s_i (non-decreasing, from t = 0)."""
    a = np.zeros(4)
    a[:len(args)] = args
    s = np.ascontiguousarray(s, dtype=np.float64)
    t = np.zeros(s.size)
    check(lib().bem2d_arclength_params(int(curve_type), _np_ptr(a), _np_ptr(s), s.size,
                                       _np_ptr(t)))
    return t


class IRReport(ctypes.Structure):
    """Outcome of the mixed-precision solve (bem2d_ir_report in include/bem2d_hip.h)."""
    _fields_ = [("iterations", _I), ("fallback_fp64", _I), ("converged", _I), ("gmres_steps", _I),
                ("backward_error", _D), ("matrix_norm_inf", _D), ("growth", _D)]


SIGNATURES = {
    "bem2d_ctx_create": (_I, [ctypes.POINTER(_P), _I]),
    "bem2d_ctx_destroy": (None, [_P]),
    "bem2d_ctx_set_stream": (_I, [_P, _P]),
    "bem2d_ctx_reset_stream": (_I, [_P]),
    "bem2d_ctx_stream": (_P, [_P]),
    "bem2d_synchronize": (_I, [_P]),
    "bem2d_ctx_release_scratch": (_I, [_P]),
    "bem2d_last_error": (ctypes.c_char_p, []),
    "bem2d_device_alloc": (_I, [_P, _I64, ctypes.POINTER(_P)]),
    "bem2d_device_free": (_I, [_P, _P]),
    "bem2d_copy_to_device": (_I, [_P, _P, _P, _I64]),
    "bem2d_copy_to_host": (_I, [_P, _P, _P, _I64]),
    "bem2d_mesh_create": (_I, [_P, ctypes.POINTER(Geometry), QuadOption, ctypes.POINTER(_P)]),
    "bem2d_mesh_destroy": (None, [_P]),
    "bem2d_mesh_size": (_I64, [_P]),
    "bem2d_mesh_special_count": (_I64, [_P]),
    "bem2d_assemble": (_I, [_P, _P, _I, _D, _D, _D, _D, _P, _I64]),
    "bem2d_rhs": (_I, [_P, _P, _I, _D, _D, _P, _I, _D, _P, _I64]),
    "bem2d_solve_fp64": (_I, [_P, _I64, _I64, _P, _I64, _P, _I64, _I]),
    "bem2d_lu32_create": (_I, [_P, _I64, ctypes.POINTER(_P)]),
    "bem2d_lu32_destroy": (None, [_P]),
    "bem2d_lu32_size": (_I64, [_P]),
    "bem2d_lu32_graph_replays": (_I64, [_P]),
    "bem2d_tune_set_assembly": (_I, [_P, _I, _P, _P]),
    "bem2d_tune_get_assembly": (_I, [_P, _P, _P]),
    "bem2d_tune_set_lu_pair_tiles": (_I, [_P, _I]),
    "bem2d_tune_get_lu_pair_tiles": (_I, [_P]),
    "bem2d_tune_set_lu_quad_tiles": (_I, [_P, _I]),
    "bem2d_tune_get_lu_quad_tiles": (_I, [_P]),
    "bem2d_tune_set_lu_oct_tiles": (_I, [_P, _I]),
    "bem2d_tune_get_lu_oct_tiles": (_I, [_P]),
    "bem2d_tune_set_lu_split_tiles": (_I, [_P, _I]),
    "bem2d_tune_get_lu_split_tiles": (_I, [_P]),
    "bem2d_tune_load": (_I, [_P, ctypes.c_char_p]),
    "bem2d_lu32_factor": (_I, [_P, _P, _P, _I64]),
    "bem2d_lu32_refine": (_I, [_P, _P, _P, _I64, _P, _I64, _P, _I64, _I, _D, _I,
                               ctypes.POINTER(IRReport)]),
    "bem2d_solve_mixed": (_I, [_P, _I64, _I64, _P, _I64, _P, _I64, _D, ctypes.POINTER(IRReport)]),
    "bem2d_solve_mixed_static": (_I, [_P, _I64, _I64, _P, _I64, _P, _I64, _D,
                                      ctypes.POINTER(IRReport)]),
    "bem2d_change_basis": (_I, [_P, _I64, _P, _I64, _P, _I64]),
    "bem2d_singular_values": (_I, [_P, _I64, _P, _I64, _P]),
    "bem2d_l2_norm_cond": (_I, [_P, _I64, _P, _I64, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "bem2d_hermitian_eigenvalues": (_I, [_P, _I64, _P, _I64, _P]),
    "bem2d_numerical_range": (_I, [_P, _I64, _P, _I64, _I, _P]),
    "bem2d_eigenvalues": (_I, [_P, _I64, _P, _I64, _P]),
    "bem2d_eval_single_layer": (_I, [_P, _P, _D, _P, _P, _I64, _P]),
    "bem2d_farfield": (_I, [_P, _P, _D, _P, _I64, _I, _P, _I64, _P]),
    "bem2d_assemble_basis": (_I, [_P, _P, _P, _I, _I, _D, _D, _D, _D, _D, _P, _I64]),
    "bem2d_rhs_basis": (_I, [_P, _P, _P, _I, _I, _D, _D, _D, _P, _I, _D, _P, _I64]),
    "bem2d_eval_single_layer_basis": (_I, [_P, _P, _P, _I, _D, _D, _P, _P, _I64, _P]),
    "bem2d_farfield_basis": (_I, [_P, _P, _P, _I, _D, _P, _I64, _I, _P, _I64, _P]),
    "bem2d_basis_values": (_I, [_P, _I, _P, _I64, _P]),
    "bem2d_curve_length": (_I, [_I, _P, _D, _D, _P]),
    "bem2d_arclength_params": (_I, [_I, _P, _P, _I64, _P]),
    "bem2d_profile_enable": (_I, [_P, _I]),
    "bem2d_profile_reset": (_I, [_P]),
    "bem2d_profile_read": (_I, [_P, _I, ctypes.POINTER(_D), ctypes.POINTER(_I64)]),
    "bem2d_gauss_legendre": (_I, [_I, _P, _P]),
    "bem2d_quad_rule": (_I, [_I, _I, _I, _D, _P, _P, _P]),
    "bem2d_hankel01": (_I, [_P, _P, _I64, _P]),
    "bem2d_hankel01_host": (_I, [_P, _I64, _P]),
    "bem2d_hankel01_complex": (_I, [_P, _D, _P, _I64, _P]),
    "bem2d_hankel01_complex_host": (_I, [_D, _P, _I64, _P]),
    "bem2d_assemble_complex": (_I, [_P, _P, _I, _D, _D, _D, _D, _D, _P, _I64]),
    "bem2d_rhs_complex": (_I, [_P, _P, _I, _D, _D, _D, _P, _I, _D, _P, _I64]),
    "bem2d_eval_single_layer_complex": (_I, [_P, _P, _D, _D, _P, _P, _I64, _P]),
}

# ---- serving engine ------------------------------------------------------------------------------
PRECISION_MIXED, PRECISION_FP64 = 0, 1


class Request(ctypes.Structure):
    """bem2d_request: one scattering problem for the serving engine (host arrays)."""
    _fields_ = [("geom", Geometry), ("opts", QuadOption), ("k", _D), ("eta", _D), ("nrhs", _I),
                ("precision", _I), ("dirs", _P), ("nang", _I64), ("theta", _P), ("tol", _D)]


class ResultInfo(ctypes.Structure):
    _fields_ = [("status", _I), ("precision", _I), ("n", _I64), ("nrhs", _I), ("iterations", _I),
                ("backward_error", _D), ("t_arrival", _D), ("t_admit", _D), ("t_done", _D),
                ("fallback_fp64", _I)]


class EngineStats(ctypes.Structure):
    _fields_ = [(name, _I64) for name in (
        "submitted", "admitted", "completed", "failed", "assembly_launches", "lu_step_launches",
        "refine_step_launches", "arena_bytes", "arena_in_use", "queued", "active")]


SIGNATURES.update({
    "bem2d_engine_footprint": (_I64, [_I64, _I, _I64, _I, QuadOption]),
    "bem2d_engine_create": (_I, [_P, _I64, _I, ctypes.POINTER(_P)]),
    "bem2d_engine_destroy": (None, [_P]),
    "bem2d_engine_now": (_D, [_P]),
    "bem2d_engine_submit": (_I, [_P, ctypes.POINTER(Request), _D, ctypes.POINTER(_I64)]),
    "bem2d_engine_step": (_I, [_P, ctypes.POINTER(_I)]),
    "bem2d_engine_poll": (_I, [_P, _I64, ctypes.POINTER(_I)]),
    "bem2d_engine_result": (_I, [_P, _I64, _P, _P, ctypes.POINTER(ResultInfo)]),
    "bem2d_engine_get_stats": (_I, [_P, ctypes.POINTER(EngineStats)]),
    "bem2d_engine_run_trace": (_I, [_P, ctypes.POINTER(Request), _P, _I, _P]),
})


_LIB = None


class Bem2dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bem2d status %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded HIP library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is not built: run `make -C bem2d_amd` (or "
                               "__graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(code):
    if code != OK:
        raise Bem2dError(code, lib().bem2d_last_error().decode(errors="replace"))


def _np_ptr(a):
    return a.ctypes.data_as(_P)


def _t_ptr(t):
    return _P(t.data_ptr())


def _ld(t):
    """Leading dimension of a column-major (stride(0) == 1) matrix view.  A single column's stride
    is never stepped, and torch may report it as 1; the C-ABI wants ld >= rows."""
    if t.stride(0) != 1 and t.shape[0] > 1:
        raise ValueError("expected a column-major view (stride(0) == 1), got strides %s"
                         % (t.stride(),))
    return max(t.stride(1), t.shape[0]) if t.shape[1] == 1 else t.stride(1)


def hankel01_host(x):
    """Host build of the device Hankel routine: returns (H0, H1) complex arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((x.size, 4))
    check(lib().bem2d_hankel01_host(_np_ptr(x), x.size, _np_ptr(out)))
    return out[:, 0] + 1j * out[:, 1], out[:, 2] + 1j * out[:, 3]


def hankel01_complex_host(phi, z):
    """Host run of the FP64 complex Hankel routine at complex z (arg z within 0.05 of phi):
    (n, 4) array of Re H0, Im H0, Re H1, Im H1."""
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros((z.size, 4))
    check(lib().bem2d_hankel01_complex_host(phi, _np_ptr(z), z.size, _np_ptr(out)))
    return out


def quad_rule(which, N, L=3, sigma=0.15):
    """Host rule of the product: Gauss2D (0), AdaptedGauss1 (1), AdaptedGauss2 (2), AdaptedGauss3
    (3) -- the tables the device assembly uses."""
    m = lib().bem2d_quad_rule(which, N, L, sigma, None, None, None)
    if m < 0:
        check(ERR_ARG)
    x, y, w = np.zeros(m), np.zeros(m), np.zeros(m)
    lib().bem2d_quad_rule(which, N, L, sigma, _np_ptr(x), _np_ptr(y), _np_ptr(w))
    return x, y, w


class Context:
    """A bem2d_ctx on one GPU, issuing work on torch's current stream of that device."""

    def __init__(self, device=0, use_torch_stream=True):
        import torch
        self.device = device
        h = _P()
        check(lib().bem2d_ctx_create(ctypes.byref(h), device))
        self.h = h
        # meshes, factorisations and engines of this context: closed before it, so that none
        # outlives the bem2d_ctx it points to (e.g. objects kept alive by a traceback)
        self._children = weakref.WeakSet()
        if use_torch_stream:
            s = torch.cuda.current_stream(device)
            check(lib().bem2d_ctx_set_stream(self.h, _P(s.cuda_stream)))

    def sync(self):
        check(lib().bem2d_synchronize(self.h))

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            lib().bem2d_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def profile(self, on=True):
        check(lib().bem2d_profile_enable(self.h, 1 if on else 0))
        check(lib().bem2d_profile_reset(self.h))

    def profile_reset(self):
        check(lib().bem2d_profile_reset(self.h))

    def release_scratch(self):
        """Frees the context's scratch buffers (they grow back on demand)."""
        check(lib().bem2d_ctx_release_scratch(self.h))

    def profile_read(self, kernel_id):
        """(total ms, launches) of an instrumented kernel since the last reset."""
        ms = _D()
        cnt = _I64()
        check(lib().bem2d_profile_read(self.h, kernel_id, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    # ---- launch-shape table ----------------------------------------------------------------
    def tune_assembly(self, buckets=None):
        """Gets (and with buckets = [(upper_tiles, split), ...] sets) the assembly table."""
        if buckets is not None:
            up = np.array([b[0] for b in buckets], dtype=np.int64)
            sp = np.array([b[1] for b in buckets], dtype=np.int32)
            check(lib().bem2d_tune_set_assembly(self.h, len(buckets), _np_ptr(up), _np_ptr(sp)))
        up = np.zeros(8, dtype=np.int64)
        sp = np.zeros(8, dtype=np.int32)
        nb = lib().bem2d_tune_get_assembly(self.h, _np_ptr(up), _np_ptr(sp))
        return [(int(up[b]), int(sp[b])) for b in range(nb)]

    def tune_lu_pair_tiles(self, tiles=None):
        if tiles is not None:
            check(lib().bem2d_tune_set_lu_pair_tiles(self.h, int(tiles)))
        return lib().bem2d_tune_get_lu_pair_tiles(self.h)

    def tune_lu_quad_tiles(self, tiles=None):
        """Trailing-matrix size (64 x 64 tiles per side) from which the FP32 LU groups four panels
        into one depth-256 trailing update; returns the current value."""
        if tiles is not None:
            check(lib().bem2d_tune_set_lu_quad_tiles(self.h, int(tiles)))
        return lib().bem2d_tune_get_lu_quad_tiles(self.h)

    def tune_lu_oct_tiles(self, tiles=None):
        """Trailing-matrix size (64 x 64 tiles per side) from which the FP32 LU groups eight panels
        into one depth-512 trailing update; returns the current value."""
        if tiles is not None:
            check(lib().bem2d_tune_set_lu_oct_tiles(self.h, int(tiles)))
        return lib().bem2d_tune_get_lu_oct_tiles(self.h)

    def tune_lu_split_tiles(self, tiles=None):
        """Trailing-matrix size (64 x 64 tiles per side) from which the FP32 LU factorises the
        diagonal block apart from the update; returns the current value."""
        if tiles is not None:
            check(lib().bem2d_tune_set_lu_split_tiles(self.h, int(tiles)))
        return lib().bem2d_tune_get_lu_split_tiles(self.h)

    def tune_load(self, path):
        check(lib().bem2d_tune_load(self.h, path.encode()))

    # ---- operations ------------------------------------------------------------------------
    def mesh(self, geom, N=16, L=3, sigma=0.15):
        return Mesh(self, geom, N, L, sigma)

    def hankel01(self, x):
        """Device Hankel: x a float64 cuda tensor; returns (n, 4) float64 tensor."""
        import torch
        out = torch.empty((x.numel(), 4), dtype=torch.float64, device=x.device)
        check(lib().bem2d_hankel01(self.h, _t_ptr(x), x.numel(), _t_ptr(out)))
        return out

    def hankel01_complex(self, phi, z):
        """Device complex Hankel: z a complex128 cuda tensor (arg z within 0.05 of phi);
        returns an (n, 4) float64 tensor."""
        import torch
        out = torch.empty((z.numel(), 4), dtype=torch.float64, device=z.device)
        check(lib().bem2d_hankel01_complex(self.h, phi, _t_ptr(z), z.numel(), _t_ptr(out)))
        return out

    def assemble(self, mesh, kind, k, eta=0.0, mass_scale=0.0, kernel_scale=1.0, out=None):
        """Returns the n x n complex128 matrix as a column-major (transposed-storage) view.
        k may be complex (freqtype {re, im}); a nonzero imaginary part takes the complex-k
        kernels."""
        import torch
        n = mesh.n
        if out is None:
            store = torch.empty((n, n), dtype=torch.complex128, device="cuda:%d" % self.device)
            out = store.t()
        assert out.stride(0) == 1 and out.shape == (n, n)
        kc = complex(k)
        if kc.imag != 0.0:
            check(lib().bem2d_assemble_complex(self.h, mesh.h, kind, kc.real, kc.imag, eta,
                                               mass_scale, kernel_scale, _t_ptr(out), _ld(out)))
        else:
            check(lib().bem2d_assemble(self.h, mesh.h, kind, kc.real, eta, mass_scale,
                                       kernel_scale, _t_ptr(out), _ld(out)))
        return out

    def rhs(self, mesh, kind, k, eta, dirs, scale=1.0, out=None):
        import torch
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 2)
        nrhs = dirs.shape[0]
        n = mesh.n
        if out is None:
            out = torch.empty((nrhs, n), dtype=torch.complex128,
                              device="cuda:%d" % self.device).t()
        kc = complex(k)
        if kc.imag != 0.0:
            check(lib().bem2d_rhs_complex(self.h, mesh.h, kind, kc.real, kc.imag, eta,
                                          _np_ptr(dirs), nrhs, scale, _t_ptr(out), _ld(out)))
        else:
            check(lib().bem2d_rhs(self.h, mesh.h, kind, kc.real, eta, _np_ptr(dirs), nrhs, scale,
                                  _t_ptr(out), _ld(out)))
        return out

    def solve_fp64(self, A, B, preserve_A=True):
        """Overwrites B (n x nrhs column-major complex128) with A^{-1} B."""
        n = A.shape[0]
        nrhs = B.shape[1]
        check(lib().bem2d_solve_fp64(self.h, n, nrhs, _t_ptr(A), _ld(A), _t_ptr(B), _ld(B),
                                     1 if preserve_A else 0))
        return B

    def solve_mixed(self, A, B, tol=1e-12):
        """Overwrites B with A^{-1} B (FP32 MFMA LU + FP64 refinement, FP64 fallback).
        Returns the IRReport."""
        rep = IRReport()
        check(lib().bem2d_solve_mixed(self.h, A.shape[0], B.shape[1], _t_ptr(A), _ld(A),
                                      _t_ptr(B), _ld(B), tol, ctypes.byref(rep)))
        return rep

    def solve_mixed_static(self, A, B, tol=1e-12, raise_on_fail=True):
        """bem2d_solve_mixed_static: the mixed solve in the library's fixed device workspace
        (n <= 1024, nrhs <= 8; no allocation).  Overwrites B on success; returns
        (status, IRReport)."""
        rep = IRReport()
        rc = lib().bem2d_solve_mixed_static(self.h, A.shape[0], B.shape[1], _t_ptr(A), _ld(A),
                                            _t_ptr(B), _ld(B), tol, ctypes.byref(rep))
        if raise_on_fail:
            check(rc)
        return rc, rep

    def lu32(self, n):
        return LU32(self, n)

    # ---- spectral drivers (device matrices: complex128, column-major views) -----------------
    def change_basis(self, K, M):
        """K <- U^{-H} K U^{-1} with M = U^H U (ChangeBasis); K is overwritten and returned."""
        check(lib().bem2d_change_basis(self.h, K.shape[0], _t_ptr(K), _ld(K), _t_ptr(M), _ld(M)))
        return K

    def singular_values(self, A):
        import torch
        s = torch.empty(A.shape[0], dtype=torch.float64, device=A.device)
        check(lib().bem2d_singular_values(self.h, A.shape[0], _t_ptr(A), _ld(A), _t_ptr(s)))
        return s

    def l2_norm_cond(self, A):
        nrm, cnd = _D(), _D()
        check(lib().bem2d_l2_norm_cond(self.h, A.shape[0], _t_ptr(A), _ld(A), ctypes.byref(nrm),
                                       ctypes.byref(cnd)))
        return nrm.value, cnd.value

    def hermitian_eigenvalues(self, A):
        import torch
        w = torch.empty(A.shape[0], dtype=torch.float64, device=A.device)
        check(lib().bem2d_hermitian_eigenvalues(self.h, A.shape[0], _t_ptr(A), _ld(A), _t_ptr(w)))
        return w

    def eigenvalues(self, A):
        """All eigenvalues of the general complex matrix A (device, unchanged): complex128 tensor."""
        import torch
        w = torch.empty(A.shape[0], dtype=torch.complex128, device=A.device)
        check(lib().bem2d_eigenvalues(self.h, A.shape[0], _t_ptr(A), _ld(A), _t_ptr(w)))
        return w

    def numerical_range(self, A, nsteps):
        """The 2 nsteps + 1 boundary points (complex, numpy) of NumRange."""
        pts = np.zeros(2 * (2 * nsteps + 1))
        check(lib().bem2d_numerical_range(self.h, A.shape[0], _t_ptr(A), _ld(A), int(nsteps),
                                          _np_ptr(pts)))
        return pts[0::2] + 1j * pts[1::2]

    def eval_single_layer(self, mesh, k, coef, pts):
        import torch
        pts = pts.contiguous()
        out = torch.empty(pts.shape[0], dtype=torch.complex128, device=pts.device)
        kc = complex(k)
        if kc.imag != 0.0:
            check(lib().bem2d_eval_single_layer_complex(self.h, mesh.h, kc.real, kc.imag,
                                                        _t_ptr(coef.contiguous()), _t_ptr(pts),
                                                        pts.shape[0], _t_ptr(out)))
        else:
            check(lib().bem2d_eval_single_layer(self.h, mesh.h, kc.real, _t_ptr(coef.contiguous()),
                                                _t_ptr(pts), pts.shape[0], _t_ptr(out)))
        return out

    def farfield(self, mesh, k, coef, theta):
        """coef: n x nrhs column-major; returns nang x nrhs column-major."""
        import torch
        nrhs = coef.shape[1]
        nang = theta.numel()
        out = torch.empty((nrhs, nang), dtype=torch.complex128, device=theta.device).t()
        check(lib().bem2d_farfield(self.h, mesh.h, k, _t_ptr(coef), _ld(coef), nrhs,
                                   _t_ptr(theta.contiguous()), nang, _t_ptr(out)))
        return out


    # ---- several bases per element (bases: list of (kind, degree, direction, k)) -------------
    def assemble_basis(self, mesh, bases, kind, k, eta=0.0, mass_scale=0.0, kernel_scale=1.0,
                       out=None):
        """The (n nb) x (n nb) matrix, row / column element * nb + basis, column-major view."""
        import torch
        nn = mesh.n * len(bases)
        if out is None:
            out = torch.empty((nn, nn), dtype=torch.complex128,
                              device="cuda:%d" % self.device).t()
        assert out.stride(0) == 1 and out.shape == (nn, nn)
        kc = complex(k)
        check(lib().bem2d_assemble_basis(self.h, mesh.h, basis_array(bases), len(bases), kind,
                                         kc.real, kc.imag, eta, mass_scale, kernel_scale,
                                         _t_ptr(out), _ld(out)))
        return out

    def rhs_basis(self, mesh, bases, kind, k, eta, dirs, scale=1.0, out=None):
        import torch
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 2)
        nrhs = dirs.shape[0]
        nn = mesh.n * len(bases)
        if out is None:
            out = torch.empty((nrhs, nn), dtype=torch.complex128,
                              device="cuda:%d" % self.device).t()
        kc = complex(k)
        check(lib().bem2d_rhs_basis(self.h, mesh.h, basis_array(bases), len(bases), kind,
                                    kc.real, kc.imag, eta, _np_ptr(dirs), nrhs, scale,
                                    _t_ptr(out), _ld(out)))
        return out

    def eval_single_layer_basis(self, mesh, bases, k, coef, pts):
        import torch
        pts = pts.contiguous()
        out = torch.empty(pts.shape[0], dtype=torch.complex128, device=pts.device)
        kc = complex(k)
        check(lib().bem2d_eval_single_layer_basis(self.h, mesh.h, basis_array(bases), len(bases),
                                                  kc.real, kc.imag, _t_ptr(coef.contiguous()),
                                                  _t_ptr(pts), pts.shape[0], _t_ptr(out)))
        return out

    def farfield_basis(self, mesh, bases, k, coef, theta):
        """coef: (n nb) x nrhs column-major; returns nang x nrhs column-major."""
        import torch
        nrhs = coef.shape[1]
        nang = theta.numel()
        out = torch.empty((nrhs, nang), dtype=torch.complex128, device=theta.device).t()
        check(lib().bem2d_farfield_basis(self.h, mesh.h, basis_array(bases), len(bases), k,
                                         _t_ptr(coef), _ld(coef), nrhs,
                                         _t_ptr(theta.contiguous()), nang, _t_ptr(out)))
        return out


class LU32:
    """FP32 factors resident on the device (bem2d_lu32)."""

    def __init__(self, ctx, n):
        self.ctx = ctx
        self.n = n
        h = _P()
        check(lib().bem2d_lu32_create(ctx.h, n, ctypes.byref(h)))
        self.h = h
        ctx._children.add(self)

    def factor(self, A):
        check(lib().bem2d_lu32_factor(self.ctx.h, self.h, _t_ptr(A), _ld(A)))

    def graph_replays(self):
        """Factorisations served by replaying the captured hipGraph."""
        return lib().bem2d_lu32_graph_replays(self.h)

    def refine(self, A, B, X, tol=1e-12, maxit=30, raise_on_fail=True):
        rep = IRReport()
        rc = lib().bem2d_lu32_refine(self.ctx.h, self.h, _t_ptr(A), _ld(A), _t_ptr(B), _ld(B),
                                     _t_ptr(X), _ld(X), B.shape[1], tol, maxit, ctypes.byref(rep))
        if raise_on_fail or rc not in (OK, ERR_NOT_CONVERGED):
            check(rc)
        return rep

    def close(self):
        if self.h:
            lib().bem2d_lu32_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """A bem2d_mesh: device copy of a Geometry plus its quadrature tables."""

    def __init__(self, ctx, geom, N=16, L=3, sigma=0.15):
        self.ctx = ctx
        self.geom = geom
        self.n = geom.n
        g = Geometry(geom.n, _np_ptr(geom.type), _np_ptr(geom.param), _np_ptr(geom.index),
                     _np_ptr(geom.next), _np_ptr(geom.prev))
        h = _P()
        check(lib().bem2d_mesh_create(ctx.h, ctypes.byref(g), QuadOption(L, N, sigma),
                                      ctypes.byref(h)))
        self.h = h
        ctx._children.add(self)
        self.N, self.L, self.sigma = N, L, sigma

    @property
    def special_count(self):
        return lib().bem2d_mesh_special_count(self.h)

    def close(self):
        if self.h:
            lib().bem2d_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """The serving engine (bem2d_engine): continuous batching of scattering requests on one GPU.

    Requests are dicts or objects with geom (Geometry), k, eta, dirs ((nrhs, 2)), theta (angles of
    the far field), precision (PRECISION_MIXED / PRECISION_FP64), tol, N, L, sigma.  Host arrays
    are copied by the engine at submission."""

    def __init__(self, ctx, hbm_bytes, max_active=64):
        self.ctx = ctx
        h = _P()
        check(lib().bem2d_engine_create(ctx.h, int(hbm_bytes), int(max_active), ctypes.byref(h)))
        self.h = h
        self._meta = {}
        ctx._children.add(self)

    @staticmethod
    def footprint(n, nrhs, nang, precision=PRECISION_MIXED, N=16, L=3, sigma=0.15):
        return lib().bem2d_engine_footprint(n, nrhs, nang, precision, QuadOption(L, N, sigma))

    def now(self):
        return lib().bem2d_engine_now(self.h)

    @staticmethod
    def _request(req, keep):
        geom = req["geom"]
        dirs = np.ascontiguousarray(req["dirs"], dtype=np.float64).reshape(-1, 2)
        theta = np.ascontiguousarray(req.get("theta", np.zeros(0)), dtype=np.float64)
        keep.append((geom, dirs, theta))
        g = Geometry(geom.n, _np_ptr(geom.type), _np_ptr(geom.param), _np_ptr(geom.index),
                     _np_ptr(geom.next), _np_ptr(geom.prev))
        return Request(g, QuadOption(req.get("L", 3), req.get("N", 16), req.get("sigma", 0.15)),
                       float(req["k"]), float(req.get("eta", req["k"])), dirs.shape[0],
                       int(req.get("precision", PRECISION_MIXED)), _np_ptr(dirs), theta.size,
                       _np_ptr(theta) if theta.size else None, float(req.get("tol", 1e-12)))

    def submit(self, req, t_arrival=-1.0):
        keep = []
        r = self._request(req, keep)
        t = _I64()
        check(lib().bem2d_engine_submit(self.h, ctypes.byref(r), t_arrival, ctypes.byref(t)))
        self._meta[t.value] = (req["geom"].n, r.nrhs, r.nang)
        return t.value

    def step(self):
        u = _I()
        check(lib().bem2d_engine_step(self.h, ctypes.byref(u)))
        return u.value

    def done(self, ticket):
        d = _I()
        check(lib().bem2d_engine_poll(self.h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def result(self, ticket, raise_on_fail=True):
        """(x (n, nrhs), far field (nang, nrhs), info) of a finished request; forgets it."""
        n, nrhs, nang = self._meta.pop(ticket)
        x = np.zeros((nrhs, n), dtype=np.complex128)
        ff = np.zeros((nrhs, max(nang, 1)), dtype=np.complex128)
        info = ResultInfo()
        rc = lib().bem2d_engine_result(self.h, ticket, _np_ptr(x), _np_ptr(ff), ctypes.byref(info))
        if raise_on_fail:
            check(rc)
        return x.T, ff[:, :nang].T, info

    def run_trace(self, reqs, arrival):
        """Replays reqs with arrival times (s, relative to the call); returns the tickets."""
        keep = []
        arr = (Request * len(reqs))(*[self._request(q, keep) for q in reqs])
        at = np.ascontiguousarray(arrival, dtype=np.float64)
        tickets = np.zeros(len(reqs), dtype=np.int64)
        check(lib().bem2d_engine_run_trace(self.h, arr, _np_ptr(at), len(reqs), _np_ptr(tickets)))
        for q, r, t in zip(reqs, arr, tickets):
            self._meta[int(t)] = (q["geom"].n, r.nrhs, r.nang)
        return [int(t) for t in tickets]

    def stats(self):
        s = EngineStats()
        check(lib().bem2d_engine_get_stats(self.h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in EngineStats._fields_}

    def close(self):
        if self.h:
            lib().bem2d_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
