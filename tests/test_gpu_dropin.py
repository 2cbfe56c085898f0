"""The C++ drop-in API (include/bem2d/, libbem2d.so) used the way the reference's drivers use
bem2d: bem2d_amd/host/scattering_driver runs SoundSoftScattering<PlaneWave, CombinedPlaneWave> on
the GPU; its matrix, solution, EvalSol value and far field are checked against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from bem2d_amd import geometry as G
from conftest import ROOT, analytic_farfield, rel_fro

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "bem2d_amd", "host", "scattering_driver")


def _run(shape, n, k, N=16, k_im=None):
    if not os.path.exists(DRIVER):
        pytest.fail("scattering_driver is not built")
    args = [DRIVER, shape, str(n), str(k), str(N)] + ([] if k_im is None else [str(k_im)])
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    d = json.loads(out.stdout)
    return {key: (np.array(v).reshape(-1, 2)[:, 0] + 1j * np.array(v).reshape(-1, 2)[:, 1])
            if isinstance(v, list) else v for key, v in d.items()}


@pytest.mark.parametrize("shape,n,k,N", [("circle", 200, 10.0, 16), ("kite", 300, 25.0, 16),
                                         ("polygon", 90, 8.0, 5)])
def test_driver_matches_oracle(oracle, ctx, shape, n, k, N):
    d = _run(shape, n, k, N)
    if shape == "circle":
        geom = G.unit_circle(n)
    elif shape == "kite":
        geom = G.kite(n)
    else:
        geom = G.polygon([(0, 0), (1.5, 0.2), (0.4, 1.1)], n // 3)
    A = oracle.assemble(geom, N, 3, 0.15, oracle.K_COMBINED, k, k, 1.0, 2.0)
    b = oracle.rhs(geom, N, oracle.F_COMBINED_PLANE, k, k, [1.0, 0.0], 2.0)
    x = oracle.lu_solve(A, b)[:, 0]
    assert rel_fro(d["diag"], np.diag(A)) < 1e-12
    assert rel_fro(d["col0"], A[:, 0]) < 1e-12
    assert rel_fro(d["solution"], x) < 1e-12
    th = 2 * np.pi * np.arange(360) / 360
    ff = oracle.farfield(geom, N, k, x, th)[:, 0]
    assert rel_fro(d["farfield"], ff) < 1e-12
    u_inc = np.exp(1j * k * -1.5)
    u = u_inc - oracle.eval_sl(geom, N, k, x, np.array([[-1.5, 0.0]]))
    assert abs(d["u_at_point"][0] - u[0]) < 1e-12 * abs(u[0])
    if shape == "circle":
        ser = analytic_farfield(k, th)
        assert np.abs(d["farfield"] - ser).max() / np.abs(ser).max() < 2.5e-3


def test_driver_complex_wavenumber(oracle, ctx):
    """SoundSoftScattering with freqtype k = {6, -0.7}: the kernels' and fields' complex branches
    through the C++ drop-in (DiscreteKernel, DotProdBasFuns, SolveSystem, EvalSol)."""
    n, kr, ki = 120, 6.0, -0.7
    d = _run("circle", n, kr, 16, ki)
    k = complex(kr, ki)
    geom = G.unit_circle(n)
    A = oracle.assemble(geom, 16, 3, 0.15, oracle.K_COMBINED, k, kr, 1.0, 2.0)
    b = oracle.rhs(geom, 16, oracle.F_COMBINED_PLANE, k, kr, [1.0, 0.0], 2.0)
    x = oracle.lu_solve(A, b)[:, 0]
    assert rel_fro(d["diag"], np.diag(A)) < 1e-12
    assert rel_fro(d["col0"], A[:, 0]) < 1e-12
    assert rel_fro(d["solution"], x) < 1e-12
    assert len(d["farfield"]) == 0
    u = np.exp(1j * k * -1.5) - oracle.eval_sl(geom, 16, k, x, np.array([[-1.5, 0.0]]))
    assert abs(d["u_at_point"][0] - u[0]) < 1e-12 * abs(u[0])
