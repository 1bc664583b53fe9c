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


# ---- learned correction (oracle_mlp.cpp) -------------------------------------

MLP_NPARAMS = 11331
MLP_IN_PAD = 128


def _mlp_lib():
    lib = load()
    if not getattr(lib, "_mlp_bound", False):
        F, D, I64, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int
        lib.orc_bf16_round.argtypes = [ctypes.c_float]
        lib.orc_bf16_round.restype = ctypes.c_float
        lib.orc_mlp_features.argtypes = [F, F, F, I64, F, I64, D, D, F]
        lib.orc_mlp_features.restype = None
        lib.orc_mlp_forward.argtypes = [F, I64, F, I, F]
        lib.orc_mlp_forward.restype = None
        lib.orc_mlp_grad.argtypes = [F, I64, F, F, I, F]
        lib.orc_mlp_grad.restype = D
        lib.orc_mlp_adam.argtypes = [F, F, F, F, I, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, I]
        lib.orc_mlp_adam.restype = None
        lib._mlp_bound = True
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def bf16_round(x):
    lib = _mlp_lib()
    return np.array([lib.orc_bf16_round(float(v)) for v in np.ravel(x)],
                    dtype=np.float32).reshape(np.shape(x))


def mlp_features(mass, vel, ncells, nodes, m_ref, v_ref):
    """Feature rows (nrows x 128, float32) of the listed nodes."""
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    nodes = np.ascontiguousarray(nodes, dtype=np.int64)
    nc = np.ascontiguousarray(ncells, dtype=np.int32)
    X = np.zeros((len(nodes), MLP_IN_PAD), dtype=np.float32)
    _mlp_lib().orc_mlp_features(_p(mass), _p(vel), _p(nc), mass.size, _p(nodes), len(nodes),
                                float(m_ref), float(v_ref), _p(X))
    return X


def mlp_forward(params, X, exact=False):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((X.shape[0], 3), dtype=np.float32)
    _mlp_lib().orc_mlp_forward(_p(params), X.shape[0], _p(X), int(exact), _p(Y))
    return Y


def mlp_grad(params, X, T, exact=False):
    """(loss, gradient as float64) of the mean squared error."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.zeros(MLP_NPARAMS, dtype=np.float64)
    loss = _mlp_lib().orc_mlp_grad(_p(params), X.shape[0], _p(X), _p(T), int(exact), _p(g))
    return loss, g


def mlp_adam(params, m, v, grad, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in place on float32 arrays."""
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    grad = np.ascontiguousarray(grad, dtype=np.float32)
    _mlp_lib().orc_mlp_adam(_p(params), _p(m), _p(v), _p(grad), params.size, lr, beta1, beta2,
                            eps, t)
