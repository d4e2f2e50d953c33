"""End-to-end parity of the GPU FP64 solve (pqdts_solve through the C ABI).

1. Trajectory: over the part of the config-1 solve that the reference
   reproduces across its own thread counts (stage 1 and the first stage-2
   Newton iteration), the GPU solve reports the same events as the reference
   binary's stdout (golden fixture): identical PCG/line-search counts, alphas
   and active-set sizes, objective to 1e-12.
2. Pi after stage 1 (the reference writes it to rank_     0_oiter     1.dat,
   pqdts.f90:2139-2165; thread-invariant in the reference to ~1e-15): the GPU's
   Pi at the end of stage 1 is within 1e-8 of it, with the identical set of
   exact zeros (the active set of the simplex projection).
3. Endpoint: at eps = 1e-13 the reference's own 1-vs-4-thread spread is 3e-9, so
   the optimum is pinned below the north-star bar: the GPU FP64 solution is
   within 1e-8 of the reference's final Pi with the same active set, for both
   Hessian-product paths; the objective agrees to 1e-12.
4. Invariants of every returned Pi: rows on the simplex.
"""
import json
import os

import numpy as np
import pytest

import pqdts_amd
import reftrace
from pqdts_amd import workloads

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("path", ["F", "G"])
def test_trajectory_prefix_matches_reference(cfg1, path):
    ref = json.load(open(os.path.join(GOLD, "cfg1_ref_prefix.json")))["events"]
    ctx = pqdts_amd.Context.from_problem(cfg1)
    ctx.set_hvp_path(path)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    n_ref = len(ref)
    seen = []

    def stop_after_prefix(c, kind, v):
        seen.append((kind, v))

    _, ev = ctx.solve(2, 200, 1e-6, 1, 0, on_event=stop_after_prefix)
    ev = [e for e in ev if e[0] not in (pqdts_amd._lib.EV_STAGE,)][:n_ref]
    assert [k for k, _ in ev] == [k for k, _ in ref]
    for (k, a), (_, b) in zip(ev, ref):
        a, b = np.array(a), np.array(b)
        if k == reftrace.EV_ACTIVE:
            assert a[0] == b[0]
        elif k == reftrace.EV_PCG:
            assert a[0] == b[0] and _rel(a[1:], b[1:]) <= 1e-9
        elif k == reftrace.EV_LS:
            assert a[0] == b[0] and a[2] == b[2] and _rel(a[1], b[1]) <= 1e-12
        elif k == reftrace.EV_ITER:
            assert tuple(a[:2]) == tuple(b[:2]) and a[7] == b[7]
            assert _rel(a[2], b[2]) <= 1e-12 and _rel(a[4], b[4]) <= 1e-9


@pytest.mark.parametrize("path", ["F", "G"])
def test_stage1_pi_matches_reference(cfg1, path):
    """Pi at the end of stage 1 (eps 1e-6, the reference's per-stage output) against
    the reference binary's: ||dPi||inf <= 1e-8 and the identical zero pattern."""
    z = np.load(os.path.join(GOLD, "cfg1_ref_states.npz"))
    Xref = z["stage1_eps1e-6"]
    ctx = pqdts_amd.Context.from_problem(cfg1)
    ctx.set_hvp_path(path)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    got = {}

    def at_stage(c, kind, v):
        if kind == pqdts_amd._lib.EV_STAGE and int(v[0]) == 1:
            got["X"] = c.get("X")

    ctx.solve(2, 200, 1e-6, 1, 0, on_event=at_stage)
    X = got["X"]
    err = float(np.max(np.abs(X - Xref)))
    print(f"stage 1 ({path}): ||dPi||inf vs reference = {err:.3e}")
    assert err <= 1e-8
    assert np.array_equal(X == 0.0, Xref == 0.0)


@pytest.mark.parametrize("path", ["F", "G", "G4"])
def test_tight_endpoint_matches_reference(cfg1, path):
    z = np.load(os.path.join(GOLD, "cfg1_ref_states.npz"))
    X1, X4 = z["final_eps1e-13_t1"], z["final_eps1e-13_t4"]
    spread = float(np.max(np.abs(X1 - X4)))
    assert spread <= 1e-8
    ctx = pqdts_amd.Context.from_problem(cfg1)
    ctx.set_hvp_path(path[0])
    if path == "G4":
        ctx.set_hvp_split(4)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    conv, ev = ctx.solve(2, 2000, 1e-13, 1, 0)
    X = ctx.get("X")
    assert conv < 1e-13
    err = float(np.max(np.abs(X - X1)))
    print(f"||dPi||inf vs reference = {err:.3e}; reference 1-vs-4-thread spread = {spread:.3e}")
    assert err <= 1e-8
    its = [v for k, v in ev if k == reftrace.EV_ITER]
    assert _rel(its[-1][2], float(z["final_eps1e-13_t1_L"])) <= 1e-12
    # same active set: the entries above 1e-9 in one are those above it in the other
    assert np.array_equal(X1 > 1e-9, X > 1e-9)
    assert np.allclose(X.sum(1), 1.0, atol=1e-12) and X.min() >= 0.0


def test_mode1_projected_gradient_runs_and_decreases(cfg1):
    ctx = pqdts_amd.Context.from_problem(cfg1)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    conv, ev = ctx.solve(1, 5, 1e-30, 1, 0)
    its = [v for k, v in ev if k == reftrace.EV_ITER]
    assert len(its) == 10  # mode 1 runs maxiter in both stages
    Ls = [v[2] for v in its]
    assert all(b <= a + 1e-15 for a, b in zip(Ls, Ls[1:]))
    X = ctx.get("X")
    assert np.allclose(X.sum(1), 1.0, atol=1e-12) and X.min() >= -1e-12


def test_mode1_matches_oracle(cfg1):
    import oracle
    ctx = pqdts_amd.Context.from_problem(cfg1)
    orc = oracle.Oracle.from_problem(cfg1)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    orc.buf("X")[:] = 1.0 / cfg1.N
    _, ev = ctx.solve(1, 3, 1e-30, 1, 0)
    _, evo = orc.solve(1, 3, 1e-30, 1, 0)
    a = [v for k, v in ev if k == reftrace.EV_ITER]
    b = [v for k, v in evo if k == reftrace.EV_ITER]
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x[7] == y[7] and _rel(x[2], y[2]) <= 1e-10


def test_start_stage2_and_maxiter_cap(cfg1):
    ctx = pqdts_amd.Context.from_problem(cfg1)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    conv, ev = ctx.solve(2, 4, 1e-30, 2, 0)
    its = [v for k, v in ev if k == reftrace.EV_ITER]
    assert [int(v[0]) for v in its] == [2, 2, 2, 2]
    stages = [v for k, v in ev if k == pqdts_amd._lib.EV_STAGE]
    assert [int(v[0]) for v in stages] == [2]


def test_device_timers_use_reference_names(cfg1):
    """With device timers on, each section's time is the device time between events
    recorded on the solver's stream: the reference's timer names appear
    (pqdts.f90:103-186 and its tstart calls), sub-sections fit inside their
    sections, and the whole minimisation fits inside the wall-clock time."""
    import time
    ctx = pqdts_amd.Context.from_problem(cfg1)
    ctx.device_timers(True)
    ctx.put("X", np.full((cfg1.M, cfg1.N), 1.0 / cfg1.N))
    t0 = time.perf_counter()
    ctx.solve(2, 200, 1e-6, 1, 0)
    wall = time.perf_counter() - t0
    t = ctx.timers()
    for name in ("min", "L", "L_O", "dL", "dL_O", "dL_dLdX", "BF", "BF_O", "CG", "active_dLdX", "Lalpha_min",
                 "doproj", "doproj2", "pdot"):
        assert name in t and t[name][0] > 0 and t[name][1] > 0.0, name
    assert t["min"][1] <= wall
    assert t["BF_O"][1] <= t["BF"][1] * (1 + 1e-6)
    assert t["dL_O"][1] + t["dL_dLdX"][1] <= t["dL"][1] * (1 + 1e-6)
    assert t["CG"][1] <= t["min"][1] and t["Lalpha_min"][1] <= t["min"][1]
    assert t["BF_O"][0] == t["BF"][0]  # one product per Hessian application on either path
