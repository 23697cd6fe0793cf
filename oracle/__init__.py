"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from cipher_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        vp = C.c_void_p
        L.orc_rhs.argtypes = [P(abi.Config), P(abi.Mesh), C.c_double, vp, vp, vp]
        L.orc_poststep.argtypes = [P(abi.Config), P(abi.Mesh), vp]
        L.orc_setup_problem.argtypes = [P(abi.Config), P(abi.Mesh), vp, vp]
        L.orc_wrms_norm.argtypes = [vp, vp, C.c_int64, C.c_double, C.c_double]
        L.orc_wrms_norm.restype = C.c_double
        L.orc_fastpow.argtypes = [C.c_double, C.c_uint]
        L.orc_fastpow.restype = C.c_double
        L.orc_sum_tseries.argtypes = [C.c_double, P(abi.TSeries)]
        L.orc_sum_tseries.restype = C.c_double
        L.orc_sum_tseries_derivative.argtypes = [C.c_double, P(abi.TSeries)]
        L.orc_sum_tseries_derivative.restype = C.c_double
        L.orc_set_intersection.argtypes = [vp, vp, vp, vp, vp]
        L.orc_set_union.argtypes = [vp, vp, vp, vp, vp]
        L.orc_invert_matrix.argtypes = [vp, vp, C.c_int]
        L.orc_interpolant.argtypes = [C.c_int, vp, vp, C.c_int]
        L.orc_interpolant_derivative.argtypes = [C.c_int, vp, vp, C.c_int]
        L.orc_simplex_projection.argtypes = [vp, vp, C.c_int]
        for nm in ("orc_sitefrac", "orc_sitepot_explicit", "orc_sitepot_implicit",
                   "orc_mobility_component"):
            getattr(L, nm).argtypes = [P(abi.Config), C.c_int, vp, vp, C.c_double]
        L.orc_sitefrac_tangent.argtypes = [P(abi.Config), C.c_int, vp, vp, vp, vp, vp, C.c_double]
        L.orc_chemenergy_phase.argtypes = [P(abi.Config), C.c_int, vp, C.c_double]
        L.orc_chemenergy_phase.restype = C.c_double
        L.orc_mobility_volumeref.argtypes = [P(abi.Config), vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def rhs(cfg, mesh, X, t=0.0, want_active=False):
    X = np.ascontiguousarray(X, dtype=np.float64)
    F = np.zeros_like(X)
    act = np.zeros(mesh.ncells, dtype=np.uint32) if want_active else None
    ms = mesh.struct()
    rc = lib().orc_rhs(C.byref(cfg.struct), C.byref(ms), t, _p(X), _p(F),
                       _p(act) if act is not None else None)
    if rc:
        raise RuntimeError(f"orc_rhs failed: {rc}")
    return (F, act) if want_active else F


def poststep(cfg, mesh, X):
    X = np.ascontiguousarray(X, dtype=np.float64).copy()
    ms = mesh.struct()
    rc = lib().orc_poststep(C.byref(cfg.struct), C.byref(ms), _p(X))
    return X, rc


def setup_problem(cfg, mesh, phase):
    D = cfg.layout()["ndof"]
    X = np.zeros(mesh.ncells * D, dtype=np.float64)
    ph = np.ascontiguousarray(phase, dtype=np.int32)
    ms = mesh.struct()
    rc = lib().orc_setup_problem(C.byref(cfg.struct), C.byref(ms), _p(ph), _p(X))
    if rc:
        raise RuntimeError(f"orc_setup_problem failed: {rc}")
    return X


def wrms_norm(U, Y, atol, rtol):
    U = np.ascontiguousarray(U, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    return lib().orc_wrms_norm(_p(U), _p(Y), U.size, atol, rtol)
