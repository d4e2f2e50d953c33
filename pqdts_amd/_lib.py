"""ctypes binding of libpqdts_amd.so (the C ABI in include/pqdts_amd.h).

The library is the product: every operator runs as HIP kernels on the GPU.
There is no CPU fallback; loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpqdts_amd.so")

SLOTS = {"X": 0, "G": 1, "HP": 2, "XTMP": 3, "SN": 4, "Z": 5}
ROW_MAJOR, COL_MAJOR = 0, 1
EV_ITER, EV_PCG, EV_LS, EV_ACTIVE, EV_STOP1, EV_ITER0, EV_LAMBDA, EV_STAGE, EV_POLISH, EV_POLISHED = range(10)
EVENT_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int)

# every symbol the header declares, for the load check
EXPORTS = [
    "pqdts_last_error", "pqdts_device_count", "pqdts_create", "pqdts_destroy", "pqdts_put", "pqdts_get",
    "pqdts_get_O", "pqdts_set_maxr", "pqdts_get_maxr", "pqdts_get_ik", "pqdts_set_fullspace",
    "pqdts_get_fullspace", "pqdts_L", "pqdts_dL", "pqdts_BF", "pqdts_Bdiag", "pqdts_doproj", "pqdts_doproj2",
    "pqdts_conv_test", "pqdts_pdot", "pqdts_paxpy", "pqdts_pgaxpy", "pqdts_pgaxpy2", "pqdts_pset",
    "pqdts_pcopy", "pqdts_pscale", "pqdts_pcopy_scale", "pqdts_transform", "pqdts_lalpha_min2",
    "pqdts_lalpha_min_auglag", "pqdts_solve", "pqdts_timers", "pqdts_simplexproj_rows",
    "pqdts_kernel_timing", "pqdts_kernel_stats", "pqdts_info", "pqdts_set_hvp_path", "pqdts_set_hvp_split",
    "pqdts_batch_create", "pqdts_batch_destroy", "pqdts_batch_admit", "pqdts_batch_step", "pqdts_batch_result",
    "pqdts_batch_release", "pqdts_batch_active", "pqdts_batch_kernel_timing", "pqdts_batch_kernel_stats",
    "pqdts_device_cu_count", "pqdts_batch_set_cu_mask", "simplexproj_condat_",
    "pqdts_simplexproj_rows_noalloc", "pqdts_server_create", "pqdts_server_destroy", "pqdts_server_submit",
    "pqdts_server_wait", "pqdts_server_result", "pqdts_server_info", "pqdts_footprint", "pqdts_device_timers",
    "pqdts_set_polish", "pqdts_polish_result", "pqdts_server_set_polish", "pqdts_set_polish_method",
    "pqdts_server_set_polish_method",
]

_lib = None

# polish methods (pqdts_set_polish_method)
POLISH_METHODS = {"cg": 0, "direct": 1}


class PqdtsError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PqdtsError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, d, i64, i32, dp = C.c_void_p, C.c_double, C.c_int64, C.c_int, C.POINTER(C.c_double)
    i32p = C.POINTER(C.c_int32)
    L.pqdts_last_error.restype = C.c_char_p
    L.pqdts_device_count.restype = i32
    L.pqdts_device_cu_count.restype = i32
    L.pqdts_batch_set_cu_mask.argtypes = [vp, C.POINTER(C.c_uint32), i32]
    L.pqdts_create.argtypes = [i64, i64, i64, i64, i32p, i32p, dp, i64, i32p, i32p, dp, d, C.POINTER(vp)]
    L.pqdts_destroy.argtypes = [vp]
    L.pqdts_put.argtypes = [vp, i32, dp, i32]
    L.pqdts_get.argtypes = [vp, i32, dp, i32]
    L.pqdts_get_O.argtypes = [vp, dp]
    L.pqdts_set_maxr.argtypes = [vp, i32p]
    L.pqdts_get_maxr.argtypes = [vp, i32p]
    L.pqdts_get_ik.argtypes = [vp, C.POINTER(C.c_int8)]
    L.pqdts_set_fullspace.argtypes = [vp, i32]
    L.pqdts_get_fullspace.argtypes = [vp, C.POINTER(i32)]
    L.pqdts_L.argtypes = [vp, i32, dp]
    L.pqdts_dL.argtypes = [vp, i32, i32]
    L.pqdts_BF.argtypes = [vp, i32, i32]
    L.pqdts_Bdiag.argtypes = [vp, dp]
    L.pqdts_doproj.argtypes = [vp, i32]
    L.pqdts_doproj2.argtypes = [vp, i32, dp]
    L.pqdts_conv_test.argtypes = [vp, i32, i32, dp]
    L.pqdts_pdot.argtypes = [vp, i32, i32, dp]
    L.pqdts_paxpy.argtypes = [vp, d, i32, i32]
    L.pqdts_pgaxpy.argtypes = [vp, d, i32, d, i32]
    L.pqdts_pgaxpy2.argtypes = [vp, d, i32, d, i32, i32]
    L.pqdts_pset.argtypes = [vp, d, i32]
    L.pqdts_pcopy.argtypes = [vp, i32, i32]
    L.pqdts_pscale.argtypes = [vp, d, i32]
    L.pqdts_pcopy_scale.argtypes = [vp, d, i32, i32]
    L.pqdts_transform.argtypes = [vp]
    L.pqdts_lalpha_min2.argtypes = [vp, dp, dp]
    L.pqdts_lalpha_min_auglag.argtypes = [vp, dp, dp]
    L.pqdts_solve.argtypes = [vp, i32, i32, d, i32, i32, EVENT_CB, vp, dp]
    L.pqdts_timers.argtypes = [vp, i32, C.c_void_p, C.POINTER(i32), dp]
    L.pqdts_simplexproj_rows.argtypes = [dp, dp, i64, i64]
    L.simplexproj_condat_.argtypes = [dp, dp, C.POINTER(C.c_int)]
    L.simplexproj_condat_.restype = None
    L.pqdts_simplexproj_rows_noalloc.argtypes = [vp, vp, i64, i64, vp]
    i64p = C.POINTER(i64)
    L.pqdts_server_create.argtypes = [i32, i64p, i32p, i32, i64, C.POINTER(vp)]
    L.pqdts_server_destroy.argtypes = [vp]
    L.pqdts_server_submit.argtypes = [vp, i64, i64, i64, i64, i32p, i32p, dp, i64, i32p, i32p, dp, d, i32, i32, d,
                                      i32, dp, i64p]
    L.pqdts_server_wait.argtypes = [vp, i64p, i32, i32]
    L.pqdts_server_result.argtypes = [vp, i64, dp, dp, C.POINTER(i32), dp]
    L.pqdts_server_info.argtypes = [vp, i64p]
    L.pqdts_footprint.argtypes = [i64, i64, i64, i64, i32p, i32p, i32, i64p, i64p]
    L.pqdts_kernel_timing.argtypes = [vp, i32]
    L.pqdts_device_timers.argtypes = [vp, i32]
    L.pqdts_kernel_stats.argtypes = [vp, i32, C.c_void_p, C.POINTER(i32), dp, dp, dp]
    L.pqdts_info.argtypes = [vp, C.POINTER(i64)]
    L.pqdts_set_hvp_path.argtypes = [vp, i32]
    L.pqdts_set_hvp_split.argtypes = [vp, i32]
    L.pqdts_set_polish.argtypes = [vp, i32, d, d, d, i32]
    L.pqdts_polish_result.argtypes = [vp, C.POINTER(i32), i64p, dp]
    L.pqdts_server_set_polish.argtypes = [vp, i32, d, d, d, i32]
    L.pqdts_set_polish_method.argtypes = [vp, i32]
    L.pqdts_server_set_polish_method.argtypes = [vp, i32]
    L.pqdts_batch_create.argtypes = [i32, C.POINTER(vp)]
    L.pqdts_batch_destroy.argtypes = [vp]
    L.pqdts_batch_admit.argtypes = [vp, vp, i32, i32, d, i32, i32, i32, EVENT_CB, vp, C.POINTER(i32)]
    L.pqdts_batch_step.argtypes = [vp, C.POINTER(i32), i32]
    L.pqdts_batch_result.argtypes = [vp, i32, dp, C.POINTER(i32)]
    L.pqdts_batch_release.argtypes = [vp, i32]
    L.pqdts_batch_active.argtypes = [vp]
    L.pqdts_batch_kernel_timing.argtypes = [vp, i32]
    L.pqdts_batch_kernel_stats.argtypes = [vp, i32, C.c_void_p, C.POINTER(i32), dp, dp, dp]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = load().pqdts_last_error().decode(errors="replace")
        raise PqdtsError(f"pqdts_amd error {rc}: {msg}")


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def device_count() -> int:
    return int(load().pqdts_device_count())


def device_cu_count() -> int:
    return int(load().pqdts_device_cu_count())


def simplexproj_rows(y: np.ndarray) -> np.ndarray:
    """Projects every row of y onto the probability simplex on the GPU."""
    y = np.ascontiguousarray(np.atleast_2d(y), dtype=np.float64)
    x = np.empty_like(y)
    _check(load().pqdts_simplexproj_rows(_p(y), _p(x), y.shape[0], y.shape[1]))
    return x


def simplexproj_condat(y: np.ndarray) -> np.ndarray:
    """One row through the reference's own entry point, simplexproj_condat_(y, x, &n)
    (condat_simplexproj.c:60): bit-identical to the reference's result."""
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    x = np.empty_like(y)
    n = C.c_int(y.size)
    load().simplexproj_condat_(_p(y), _p(x), C.byref(n))
    return x


def cu_mask_words(cus, ncu: int = 0):
    """(uint32 array, word count) with bit c set for every CU c in `cus`, covering at
    least ncu CUs, so that every CU of the device has an explicit bit."""
    cus = list(cus)
    n = max(max(cus) + 1 if cus else 1, ncu) // 32 + 1
    words = (C.c_uint32 * n)()
    for c in cus:
        words[c // 32] |= 1 << (c % 32)
    return words, n


def _slot(s):
    return SLOTS[s] if isinstance(s, str) else int(s)


class Context:
    """A device-resident problem: F (D x M) and P (D x N) as COO, gamma."""

    def __init__(self, M, N, D, F_coo, P_coo, gamma):
        L = load()
        fr = np.ascontiguousarray(F_coo[0], np.int32)
        fc = np.ascontiguousarray(F_coo[1], np.int32)
        fv = np.ascontiguousarray(F_coo[2], np.float64)
        pr = np.ascontiguousarray(P_coo[0], np.int32)
        pc = np.ascontiguousarray(P_coo[1], np.int32)
        pv = np.ascontiguousarray(P_coo[2], np.float64)
        self.M, self.N, self.D = int(M), int(N), int(D)
        h = C.c_void_p()
        _check(L.pqdts_create(self.M, self.N, self.D, fv.size, _p(fr, C.c_int32), _p(fc, C.c_int32), _p(fv),
                              pv.size, _p(pr, C.c_int32), _p(pc, C.c_int32), _p(pv), float(gamma), C.byref(h)))
        self.h = h

    @classmethod
    def from_problem(cls, prob):
        import scipy.sparse as sps
        F = sps.coo_matrix(prob.F)
        P = sps.coo_matrix(prob.P)
        return cls(prob.M, prob.N, prob.D, (F.row, F.col, F.data), (P.row, P.col, P.data), prob.gamma)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().pqdts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put(self, slot, arr, layout=ROW_MAJOR):
        a = np.ascontiguousarray(arr, dtype=np.float64) if layout == ROW_MAJOR else np.asfortranarray(arr, np.float64)
        assert a.size == self.M * self.N
        _check(load().pqdts_put(self.h, _slot(slot), _p(a), layout))

    def get(self, slot, layout=ROW_MAJOR) -> np.ndarray:
        out = np.empty((self.M, self.N), dtype=np.float64, order="C" if layout == ROW_MAJOR else "F")
        _check(load().pqdts_get(self.h, _slot(slot), _p(out), layout))
        return out

    def get_O(self):
        out = np.empty((self.D, self.N))
        _check(load().pqdts_get_O(self.h, _p(out)))
        return out

    @property
    def maxr(self):
        out = np.empty(self.M, np.int32)
        _check(load().pqdts_get_maxr(self.h, _p(out, C.c_int32)))
        return out

    @maxr.setter
    def maxr(self, v):
        a = np.ascontiguousarray(v, np.int32)
        _check(load().pqdts_set_maxr(self.h, _p(a, C.c_int32)))

    @property
    def ik(self):
        out = np.empty((self.M, self.N), np.int8)
        _check(load().pqdts_get_ik(self.h, _p(out, C.c_int8)))
        return out

    @property
    def fullspace(self) -> bool:
        v = C.c_int()
        _check(load().pqdts_get_fullspace(self.h, C.byref(v)))
        return bool(v.value)

    @fullspace.setter
    def fullspace(self, full: bool):
        _check(load().pqdts_set_fullspace(self.h, int(bool(full))))

    def _scalar(self, fn, *args):
        v = C.c_double()
        _check(fn(self.h, *args, C.byref(v)))
        return v.value

    def L(self, x="X"):
        return self._scalar(load().pqdts_L, _slot(x))

    def dL(self, x="X", g="G"):
        _check(load().pqdts_dL(self.h, _slot(x), _slot(g)))

    def BF(self, dd="XTMP", bx="HP"):
        _check(load().pqdts_BF(self.h, _slot(dd), _slot(bx)))

    def Bdiag(self):
        out = np.empty(self.M)
        _check(load().pqdts_Bdiag(self.h, _p(out)))
        return out

    def doproj(self, x="X"):
        _check(load().pqdts_doproj(self.h, _slot(x)))

    def doproj2(self, x="X"):
        return self._scalar(load().pqdts_doproj2, _slot(x))

    def conv_test(self, x="X", g="G"):
        return self._scalar(load().pqdts_conv_test, _slot(x), _slot(g))

    def pdot(self, x, y):
        return self._scalar(load().pqdts_pdot, _slot(x), _slot(y))

    def paxpy(self, a, x, y):
        _check(load().pqdts_paxpy(self.h, float(a), _slot(x), _slot(y)))

    def pgaxpy(self, a, x, b, y):
        _check(load().pqdts_pgaxpy(self.h, float(a), _slot(x), float(b), _slot(y)))

    def pgaxpy2(self, a, x, b, y, z):
        _check(load().pqdts_pgaxpy2(self.h, float(a), _slot(x), float(b), _slot(y), _slot(z)))

    def pset(self, a, x):
        _check(load().pqdts_pset(self.h, float(a), _slot(x)))

    def pcopy(self, x, y):
        _check(load().pqdts_pcopy(self.h, _slot(x), _slot(y)))

    def pscale(self, a, x):
        _check(load().pqdts_pscale(self.h, float(a), _slot(x)))

    def pcopy_scale(self, a, x, y):
        _check(load().pqdts_pcopy_scale(self.h, float(a), _slot(x), _slot(y)))

    def transform(self):
        _check(load().pqdts_transform(self.h))

    def lalpha_min2(self):
        a, dl = C.c_double(), C.c_double()
        _check(load().pqdts_lalpha_min2(self.h, C.byref(a), C.byref(dl)))
        return a.value, dl.value

    def lalpha_min_auglag(self):
        a, dl = C.c_double(), C.c_double()
        _check(load().pqdts_lalpha_min_auglag(self.h, C.byref(a), C.byref(dl)))
        return a.value, dl.value

    def solve(self, mode=2, maxiter=200, eps=1e-6, start_stage=1, output=0, on_event=None):
        """Runs the minimisation from the current X; returns (conv, events)."""
        events = []

        def cb(_u, kind, v, nv):
            ev = (kind, tuple(v[i] for i in range(nv)))
            events.append(ev)
            if on_event is not None:
                on_event(self, *ev)

        fn = EVENT_CB(cb)
        conv = C.c_double()
        _check(load().pqdts_solve(self.h, int(mode), int(maxiter), float(eps), int(start_stage), int(output), fn,
                                  None, C.byref(conv)))
        return conv.value, events

    def info(self) -> dict:
        a = (C.c_int64 * 14)()
        _check(load().pqdts_info(self.h, a))
        return {"M": a[0], "N": a[1], "D": a[2], "nspan": a[3], "tile_elems": a[4], "row_tiles": a[5],
                "col_tiles": a[6], "gram": bool(a[7]), "gram_tile_elems": a[8], "gram_nnz": a[9],
                "ft_tile_elems": a[10], "device_bytes": a[11], "hvp_split": a[12], "gram_nch": a[13]}

    def set_hvp_path(self, path):
        """0 automatic, 1 the two F products, 2 the Gram matrix G = F^T F."""
        _check(load().pqdts_set_hvp_path(self.h, {"auto": 0, "F": 1, "G": 2}.get(path, path)))

    def set_hvp_split(self, split: int):
        """Split of the Gram product over band chunks: 0 by shape, 1 none, 2..16 fixed."""
        _check(load().pqdts_set_hvp_split(self.h, int(split)))

    def set_polish(self, max_outer=100, cg_rtol=1e-6, tol=1e-10, budget=0.0, fp32=False, method=0):
        """Polish every later solve of this context to the problem's minimiser (primal-dual
        active-set Newton after conv < eps) to about `tol`; max_outer = 0 turns it off.
        method 0: Newton-CG steps, with at most `budget` times the minimisation's PCG rounds
        (0: no limit); method 1 ("direct"): direct face solves."""
        _check(load().pqdts_set_polish(self.h, int(max_outer), float(cg_rtol), float(tol), float(budget),
                                       int(bool(fp32))))
        _check(load().pqdts_set_polish_method(self.h, POLISH_METHODS.get(method, method)))

    def polish_result(self) -> dict:
        st = (C.c_int * 3)()
        nact = C.c_int64()
        step = C.c_double()
        _check(load().pqdts_polish_result(self.h, st, C.byref(nact), C.byref(step)))
        return {"status": st[0], "iterations": st[1], "cg_rounds": st[2], "nact": nact.value, "step": step.value}

    def kernel_timing(self, on):
        """True / 1: time every launch; k > 1: every k-th launch; False / 0: off."""
        _check(load().pqdts_kernel_timing(self.h, int(on)))

    def kernel_stats(self) -> dict:
        """{kernel: (launches, total_ms, algorithmic flops, algorithmic bytes)} from HIP events on the
        context's stream."""
        return _kstats(load().pqdts_kernel_stats, self.h)

    def device_timers(self, on: bool):
        """Section timers measure device time (HIP events on the context's stream)."""
        _check(load().pqdts_device_timers(self.h, int(bool(on))))

    def timers(self):
        L = load()
        n = L.pqdts_timers(self.h, 0, None, None, None)
        if n <= 0:
            return {}
        names = (C.c_char * 24 * n)()
        counts = (C.c_int * n)()
        secs = (C.c_double * n)()
        L.pqdts_timers(self.h, n, C.cast(names, C.c_void_p), counts, secs)
        return {bytes(names[i]).split(b"\0")[0].decode(): (counts[i], secs[i]) for i in range(n)}


def _kstats(fn, h):
    n = fn(h, 0, None, None, None, None, None)
    if n <= 0:
        return {}
    names = (C.c_char * 24 * n)()
    counts = (C.c_int * n)()
    ms = (C.c_double * n)()
    fl = (C.c_double * n)()
    by = (C.c_double * n)()
    fn(h, n, C.cast(names, C.c_void_p), counts, ms, fl, by)
    return {bytes(names[i]).split(b"\0")[0].decode(): (counts[i], ms[i], fl[i], by[i]) for i in range(n)}


class Batch:
    """Many problems advanced together (pqdts_batch_*): one launch per operator
    over all admitted problems, one host synchronisation per step."""

    def __init__(self, capacity: int):
        h = C.c_void_p()
        _check(load().pqdts_batch_create(int(capacity), C.byref(h)))
        self.h = h
        self.capacity = int(capacity)
        self.slots = {}   # slot -> (ctx, events list, callback keep-alive)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().pqdts_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, ctx: "Context", mode=2, maxiter=200, eps=1e-6, start_stage=1, output=0, precision=0,
              record=True) -> int:
        events = []

        def cb(_u, kind, v, nv):
            if record:
                events.append((kind, tuple(v[i] for i in range(nv))))

        fn = EVENT_CB(cb)
        slot = C.c_int()
        _check(load().pqdts_batch_admit(self.h, ctx.h, int(mode), int(maxiter), float(eps), int(start_stage),
                                        int(output), int(precision), fn, None, C.byref(slot)))
        self.slots[slot.value] = (ctx, events, fn)
        return slot.value

    def step(self):
        """One Newton iteration of every unfinished problem; returns finished slots."""
        buf = (C.c_int * max(1, self.capacity))()
        n = load().pqdts_batch_step(self.h, buf, self.capacity)
        if n < 0:
            _check(-1)
        return [buf[i] for i in range(min(n, self.capacity))]

    def result(self, slot):
        conv = C.c_double()
        stats = (C.c_int * 4)()
        _check(load().pqdts_batch_result(self.h, int(slot), C.byref(conv), stats))
        return {"conv": conv.value, "finished": bool(stats[0]), "newton": stats[1], "pcg": stats[2],
                "trials": stats[3]}

    def events(self, slot):
        return self.slots[slot][1]

    def release(self, slot):
        _check(load().pqdts_batch_release(self.h, int(slot)))
        return self.slots.pop(slot, None)

    def active(self) -> int:
        return int(load().pqdts_batch_active(self.h))

    def set_cus(self, cus):
        """Restricts this (empty) batch to the compute units listed in `cus`, which
        must form one consecutive range (the library refuses other masks)."""
        _check(load().pqdts_batch_set_cu_mask(self.h, *cu_mask_words(cus, device_cu_count())))

    def kernel_timing(self, on):
        """True / 1: time every launch; k > 1: every k-th launch; False / 0: off."""
        _check(load().pqdts_batch_kernel_timing(self.h, int(on)))

    def kernel_stats(self) -> dict:
        return _kstats(load().pqdts_batch_kernel_stats, self.h)

    def run(self):
        """Steps until every admitted problem has finished; returns {slot: result}."""
        done = {}
        while len(done) < len(self.slots):
            for s in self.step():
                if s not in done:
                    done[s] = self.result(s)
        return done


def _coo(A):
    import scipy.sparse as sps
    c = sps.coo_matrix(A)
    return (np.ascontiguousarray(c.row, np.int32), np.ascontiguousarray(c.col, np.int32),
            np.ascontiguousarray(c.data, np.float64))


def coo_of(prob):
    """(F, P) of a problem as the int32 / float64 COO triples the C ABI takes."""
    return _coo(prob.F), _coo(prob.P)


def footprint(prob, precision=0):
    """(persistent, transient) device bytes the problem will hold (pqdts_footprint)."""
    fr, fc, _ = _coo(prob.F)
    a, b = C.c_int64(), C.c_int64()
    _check(load().pqdts_footprint(prob.M, prob.N, prob.D, fr.size, _p(fr, C.c_int32), _p(fc, C.c_int32),
                                  int(precision), C.byref(a), C.byref(b)))
    return a.value, b.value


class Server:
    """The library's request server (pqdts_server_*): size-class lanes with continuous
    batching and admission against a device-byte budget."""

    def __init__(self, lane_bounds=(), lane_cus=None, capacity=256, budget_bytes=0):
        nl = len(lane_bounds) + 1
        b = (C.c_int64 * max(1, len(lane_bounds)))(*lane_bounds)
        cus = None
        if lane_cus is not None:
            flat = []
            for r in lane_cus:
                flat += [-1, -1] if r is None else [int(r[0]), int(r[-1])]
            cus = (C.c_int32 * len(flat))(*flat)
        h = C.c_void_p()
        _check(load().pqdts_server_create(nl, b, cus, int(capacity), int(budget_bytes), C.byref(h)))
        self.h = h
        self.shape = {}

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().pqdts_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, prob, mode=2, maxiter=5000, eps=None, precision=0, X0=None, coo=None) -> int:
        """coo: (F COO, P COO) as prepared by coo_of(prob), to keep the conversion out of
        the submission path."""
        (fr, fc, fv), (pr, pc, pv) = coo if coo is not None else coo_of(prob)
        x0 = None if X0 is None else np.ascontiguousarray(X0, np.float64)
        rid = C.c_int64()
        _check(load().pqdts_server_submit(
            self.h, prob.M, prob.N, prob.D, fv.size, _p(fr, C.c_int32), _p(fc, C.c_int32), _p(fv), pv.size,
            _p(pr, C.c_int32), _p(pc, C.c_int32), _p(pv), float(prob.gamma), int(mode), int(maxiter),
            float(prob.eps if eps is None else eps), int(precision), None if x0 is None else _p(x0),
            C.byref(rid)))
        self.shape[rid.value] = (prob.M, prob.N)
        return rid.value

    def wait(self, timeout_ms=-1, cap=4096):
        buf = (C.c_int64 * cap)()
        n = load().pqdts_server_wait(self.h, buf, cap, int(timeout_ms))
        if n < 0:
            _check(-1)
        return [buf[i] for i in range(n)]

    def result(self, rid):
        M, N = self.shape.pop(rid)
        X = np.empty((M, N))
        conv = C.c_double()
        stats = (C.c_int * 7)()
        times = (C.c_double * 3)()
        _check(load().pqdts_server_result(self.h, int(rid), _p(X), C.byref(conv), stats, times))
        return {"X": X, "conv": conv.value, "finished": bool(stats[0]), "newton": stats[1], "pcg": stats[2],
                "trials": stats[3], "queued_s": times[0], "run_s": times[1], "latency_s": times[2],
                "polish": {"status": stats[4], "iterations": stats[5], "cg_rounds": stats[6]}}

    def set_polish(self, max_outer=100, cg_rtol=1e-6, tol=1e-10, budget=0.0, fp32=False, method=0):
        """Polish settings (Context.set_polish) for requests submitted from now on."""
        _check(load().pqdts_server_set_polish(self.h, int(max_outer), float(cg_rtol), float(tol), float(budget),
                                              int(bool(fp32))))
        _check(load().pqdts_server_set_polish_method(self.h, POLISH_METHODS.get(method, method)))

    def info(self) -> dict:
        a = (C.c_int64 * 8)()
        _check(load().pqdts_server_info(self.h, a))
        keys = ("queued", "running", "finished_unclaimed", "bytes_in_use", "peak_bytes", "budget", "submitted",
                "completed")
        return dict(zip(keys, list(a)))
