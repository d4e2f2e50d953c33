"""ctypes wrapper over the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke entry point and the CPU-baseline leg of the benchmark
import this package; the product path (pqdts_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_BIN = os.path.join(REF_DIR, "pqdts_omp.x")
REF_CONDAT = os.path.join(REF_DIR, "libcondat_ref.so")

SLOTS = {"X": 0, "G": 1, "HP": 2, "XTMP": 3, "SN": 4, "Z": 5}
EV_ITER, EV_PCG, EV_LS, EV_ACTIVE, EV_STOP1, EV_ITER0, EV_LAMBDA = range(7)
EVENT_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def build_ref() -> bool:
    """Compiles the upstream sources into oracle/_ref when they are present."""
    if not os.path.isdir("/root/reference"):
        return False
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile.ref")], check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i64, i32p, f64p, vp = C.c_double, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i64, i64, i64, i64, i32p, i32p, f64p, i64, i32p, i32p, f64p, d]
        L.orc_destroy.argtypes = [vp]
        L.orc_buf.restype = f64p
        L.orc_buf.argtypes = [vp, C.c_int]
        L.orc_O.restype = f64p
        L.orc_O.argtypes = [vp]
        L.orc_diag.restype = f64p
        L.orc_diag.argtypes = [vp]
        L.orc_maxr.restype = i32p
        L.orc_maxr.argtypes = [vp]
        L.orc_ik.restype = C.POINTER(C.c_int8)
        L.orc_ik.argtypes = [vp]
        L.orc_set_fullspace.argtypes = [vp, C.c_int]
        L.orc_get_fullspace.argtypes = [vp]
        L.orc_set_mode.argtypes = [vp, C.c_int]
        for name in ("orc_redO", "orc_doproj", "orc_Bdiag", "orc_transform"):
            getattr(L, name).argtypes = [vp] + ([C.c_int] if name in ("orc_redO", "orc_doproj") else [])
        L.orc_L.restype = d
        L.orc_L.argtypes = [vp, C.c_int]
        L.orc_dL.argtypes = [vp, C.c_int, C.c_int]
        L.orc_BF.argtypes = [vp, C.c_int, C.c_int]
        L.orc_doproj2.restype = d
        L.orc_doproj2.argtypes = [vp, C.c_int]
        L.orc_conv_test.restype = d
        L.orc_conv_test.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.orc_pdot.restype = d
        L.orc_pdot.argtypes = [vp, C.c_int, C.c_int]
        L.orc_lalpha_min2.argtypes = [vp, f64p, f64p]
        L.orc_lalpha_min_auglag.argtypes = [vp, f64p, f64p]
        L.orc_simplexproj_condat.argtypes = [f64p, f64p, C.c_int]
        L.orc_simplexproj_sort.argtypes = [f64p, f64p, C.c_int]
        L.orc_solve.restype = d
        L.orc_solve.argtypes = [vp, C.c_int, C.c_int, d, C.c_int, C.c_int, EVENT_CB, vp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def condat(y: np.ndarray) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.empty_like(y)
    lib().orc_simplexproj_condat(_p(y, C.c_double), _p(x, C.c_double), y.size)
    return x


def sortproj(y: np.ndarray) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.empty_like(y)
    lib().orc_simplexproj_sort(_p(y, C.c_double), _p(x, C.c_double), y.size)
    return x


def ref_condat(y: np.ndarray) -> np.ndarray:
    """The reference's own condat_simplexproj.c, compiled into oracle/_ref."""
    L = C.CDLL(REF_CONDAT)
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.empty_like(y)
    n = C.c_int(y.size)
    L.simplexproj_condat_(_p(y, C.c_double), _p(x, C.c_double), C.byref(n))
    return x


class Oracle:
    """One problem instance held by the oracle (row-major M x N buffers)."""

    def __init__(self, M, N, D, F_coo, P_coo, gamma):
        fr, fc, fv = (np.ascontiguousarray(F_coo[0], np.int32), np.ascontiguousarray(F_coo[1], np.int32),
                      np.ascontiguousarray(F_coo[2], np.float64))
        pr, pc, pv = (np.ascontiguousarray(P_coo[0], np.int32), np.ascontiguousarray(P_coo[1], np.int32),
                      np.ascontiguousarray(P_coo[2], np.float64))
        self.M, self.N, self.D = int(M), int(N), int(D)
        self._keep = (fr, fc, fv, pr, pc, pv)
        self.h = lib().orc_create(M, N, D, fv.size, _p(fr, C.c_int32), _p(fc, C.c_int32), _p(fv, C.c_double),
                                  pv.size, _p(pr, C.c_int32), _p(pc, C.c_int32), _p(pv, C.c_double), float(gamma))

    @classmethod
    def from_problem(cls, prob):
        import scipy.sparse as sps
        F = sps.coo_matrix(prob.F)
        P = sps.coo_matrix(prob.P)
        return cls(prob.M, prob.N, prob.D, (F.row, F.col, F.data), (P.row, P.col, P.data), prob.gamma)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def buf(self, slot) -> np.ndarray:
        s = SLOTS[slot] if isinstance(slot, str) else slot
        p = lib().orc_buf(self.h, s)
        return np.ctypeslib.as_array(p, shape=(self.M, self.N))

    @property
    def O(self):
        return np.ctypeslib.as_array(lib().orc_O(self.h), shape=(self.D, self.N))

    @property
    def diag(self):
        return np.ctypeslib.as_array(lib().orc_diag(self.h), shape=(self.M,))

    @property
    def maxr(self):
        return np.ctypeslib.as_array(lib().orc_maxr(self.h), shape=(self.M,))

    @property
    def ik(self):
        return np.ctypeslib.as_array(lib().orc_ik(self.h), shape=(self.M, self.N))

    def set_fullspace(self, full: bool):
        lib().orc_set_fullspace(self.h, int(full))

    @property
    def fullspace(self) -> bool:
        return bool(lib().orc_get_fullspace(self.h))

    def L(self, slot="X"):
        return lib().orc_L(self.h, SLOTS[slot])

    def dL(self, x="X", g="G"):
        lib().orc_dL(self.h, SLOTS[x], SLOTS[g])

    def BF(self, dd="XTMP", bx="HP"):
        lib().orc_BF(self.h, SLOTS[dd], SLOTS[bx])

    def Bdiag(self):
        lib().orc_Bdiag(self.h)

    def redO(self, x="X"):
        lib().orc_redO(self.h, SLOTS[x])

    def doproj(self, x="X"):
        lib().orc_doproj(self.h, SLOTS[x])

    def doproj2(self, x="X"):
        return lib().orc_doproj2(self.h, SLOTS[x])

    def conv_test(self, x="X", g="G", prt=False):
        return lib().orc_conv_test(self.h, SLOTS[x], SLOTS[g], int(prt))

    def pdot(self, x, y):
        return lib().orc_pdot(self.h, SLOTS[x], SLOTS[y])

    def transform(self):
        lib().orc_transform(self.h)

    def lalpha_min2(self):
        a, dl = C.c_double(), C.c_double()
        lib().orc_lalpha_min2(self.h, C.byref(a), C.byref(dl))
        return a.value, dl.value

    def lalpha_min_auglag(self):
        a, dl = C.c_double(), C.c_double()
        lib().orc_lalpha_min_auglag(self.h, C.byref(a), C.byref(dl))
        return a.value, dl.value

    def solve(self, mode=2, maxiter=200, eps=1e-6, start_stage=1, output=0):
        events = []

        def cb(_u, kind, v, nv):
            events.append((kind, tuple(v[i] for i in range(nv))))

        fn = EVENT_CB(cb)
        conv = lib().orc_solve(self.h, mode, maxiter, eps, start_stage, output, fn, None)
        return conv, events
