"""The product's FP64 Hankel routine (bem2d_amd/csrc/hankel.h), run through its host build
bem2d_hankel01_host, against mpmath at 50 digits.  The same source runs on the device."""
import numpy as np

from bem2d_amd import capi
from conftest import golden


def test_host_hankel_matches_mpmath():
    g = golden("hankel_mpmath.npz")
    x, v = g["x"], g["j0y0j1y1"]
    H0, H1 = capi.hankel01_host(x)
    r0 = v[:, 0] + 1j * v[:, 1]
    r1 = v[:, 2] + 1j * v[:, 3]
    assert (np.abs(H0 - r0) / np.abs(r0)).max() < 1e-15
    assert (np.abs(H1 - r1) / np.abs(r1)).max() < 1e-15


def test_host_hankel_piece_boundaries_continuous():
    for b in (2.0, 2.6, 3.5, 5.0, 8.0, 40.0):
        x = np.array([np.nextafter(b, 0), b])
        H0, H1 = capi.hankel01_host(x)
        assert abs(H0[1] - H0[0]) < 2e-15 and abs(H1[1] - H1[0]) < 2e-15
