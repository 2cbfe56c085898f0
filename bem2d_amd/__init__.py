"""bem2d_amd: the bem2d dense Helmholtz BEM hot path on MI355X (gfx950).

The product is the C-ABI library libbem2d_hip.so (include/bem2d_hip.h) and the C++ drop-in
headers under include/bem2d/.  This Python package binds the C-ABI for the tests and benchmarks.
"""
from . import capi, geometry  # noqa: F401
from .capi import Bem2dError, Context, Engine, Mesh  # noqa: F401
