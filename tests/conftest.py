import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem2d_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture
def progress(request):
    """A print that reaches the terminal while pytest captures the test's output: the tests that
    spend minutes on the host (the complete oracle matrix of a large case) report each stage, so a
    run is never silent for long."""
    capman = request.config.pluginmanager.getplugin("capturemanager")

    def say(msg):
        if capman is None:
            print(msg, flush=True)
            return
        with capman.global_and_fixture_disabled():
            print(msg, flush=True)

    return say


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def analytic_farfield(k, theta, radius=1.0, nextra=40):
    """Far field of the sound-soft circle for the plane wave exp(ikx):
    u_inf(theta) = -sqrt(2/(pi k)) e^{-i pi/4} sum_n J_n(ka)/H_n(ka) e^{in theta}."""
    import scipy.special as sp
    nmax = int(k * radius + nextra)
    ns = np.arange(-nmax, nmax + 1)
    coef = sp.jv(ns, k * radius) / sp.hankel1(ns, k * radius)
    return -np.sqrt(2 / (np.pi * k)) * np.exp(-1j * np.pi / 4) * (
        coef[:, None] * np.exp(1j * ns[:, None] * np.asarray(theta)[None, :])).sum(axis=0)


def rel_fro(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))
