"""CPU oracle of ExaGOOP's explicit MPM step — TEST INFRASTRUCTURE ONLY.

The parity checker and the timed CPU baseline.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package; the product (exagoop_amd) never does.

liboracle_mpm.so restates the reference step (see oracle_mpm.cpp for the
file:line of each function).  Pins (tests/test_oracle_pins.py) check it
against closed forms, invariants and brute force.  AMReX-provided pieces
(getParticleCell, Geometry dx, Redistribute's periodic wrap and removal,
FillBoundary, neighbour image selection) are "parity unpinned": AMReX is
absent from the reference tree.
"""
import ctypes
import os

import numpy as np

from exagoop_amd.capi import Specs, SIGNATURES, _Base, dptr, UDF_FN  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_mpm.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        oname = "orc_" + name[4:]
        if not hasattr(lib, oname):
            continue
        fn = getattr(lib, oname)
        fn.restype = res
        fn.argtypes = args
    lib.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_threads.restype = None
    lib.orc_fill_neighbors.argtypes = [ctypes.c_void_p]
    lib.orc_fill_neighbors.restype = ctypes.c_int
    for nm in ("orc_quadraticspline_1d", "orc_quadraticspline_1d_der", "orc_cubicspline_1d",
               "orc_cubicspline_1d_der"):
        getattr(lib, nm).argtypes = [ctypes.c_int, ctypes.c_double]
        getattr(lib, nm).restype = ctypes.c_double
    i3 = ctypes.c_int * 3
    d3 = ctypes.c_double * 3
    lib.orc_basisval_and_grad.argtypes = [i3, i3, d3, d3, d3, i3, i3, i3, i3, d3]
    lib.orc_basisval_and_grad.restype = ctypes.c_double
    lib.orc_applybc.argtypes = [d3, d3, ctypes.c_double, d3, ctypes.c_int]
    lib.orc_applybc.restype = ctypes.c_int
    _lib = lib
    return lib


class Oracle(_Base):
    """CPU restatement context; same methods as exagoop_amd.Context."""

    prefix = "orc_"

    def __init__(self, specs, threads=1):
        self._lib = load()
        self.specs = specs.copy()
        ctx = ctypes.c_void_p()
        self._ctx = None
        self._call("create", ctypes.byref(specs), ctypes.byref(ctx))
        self._ctx = ctx
        n = specs.ncells
        self.nn = (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
        self._lib.orc_set_threads(self._ctx, threads)

    def set_threads(self, n):
        self._lib.orc_set_threads(self._ctx, n)

    def fill_neighbors(self):
        self._call("fill_neighbors", self._ctx)


def spline(kind, t, z, der=False):
    lib = load()
    name = {("quad", False): "orc_quadraticspline_1d", ("quad", True): "orc_quadraticspline_1d_der",
            ("cubic", False): "orc_cubicspline_1d", ("cubic", True): "orc_cubicspline_1d_der"}[
        (kind, der)]
    return getattr(lib, name)(t, z)


def basisval_and_grad(stencil, cell, xp, plo, dx, order, periodic, lo, hi):
    lib = load()
    i3 = ctypes.c_int * 3
    d3 = ctypes.c_double * 3
    g = d3()
    N = lib.orc_basisval_and_grad(i3(*stencil), i3(*cell), d3(*xp), d3(*plo), d3(*dx),
                                  i3(*order), i3(*periodic), i3(*lo), i3(*hi), g)
    return N, np.array(g[:])


def applybc(relvel, mu, normal, bc, preload=None):
    lib = load()
    d3 = ctypes.c_double * 3
    out = d3(*(relvel if preload is None else preload))
    rc = lib.orc_applybc(d3(*relvel), out, mu, d3(*normal), bc)
    return rc, np.array(out[:])
