"""Batched / served form: many uneven problems advanced together.

* Each problem's FP64 result is bit-identical alone (pqdts_solve) and in a batch
  (same per-problem arithmetic order; the batch only changes which blocks run
  together), including problems admitted while others are mid-solve.
* Mixed precision (FP32 Hessian products inside CG, FP64 elsewhere) lands within
  ||dPi||_inf <= 1e-6 of the FP64 solution when both are converged tightly, and
  the relative objective gap is reported.
"""
import numpy as np
import pytest

import pqdts_amd
from pqdts_amd import workloads

pytestmark = pytest.mark.gpu


def problems():
    return [
        workloads.config1(),
        workloads.random_problem(40, 15, 80, 0.7, 3e-3, 1e6, seed=11, eps=1e-7, name="a"),
        workloads.random_problem(77, 131, 154, 0.9, 1e-5, 1e7, seed=12, eps=1e-7, name="b"),
        workloads.random_problem(130, 60, 260, 0.6, 1e-2, 3e6, seed=13, eps=1e-7, name="c"),
    ]


def start(ctx, prob):
    ctx.put("X", np.full((prob.M, prob.N), 1.0 / prob.N))


def solve_alone(prob, maxiter=300, setup=None):
    ctx = pqdts_amd.Context.from_problem(prob)
    if setup:
        setup(ctx)
    start(ctx, prob)
    conv, ev = ctx.solve(2, maxiter, prob.eps, 1, 0)
    return ctx.get("X"), conv, [e for e in ev]


def test_batch_equals_alone_bitwise():
    probs = problems()
    alone = [solve_alone(p) for p in probs]
    b = pqdts_amd.Batch(8)
    ctxs = []
    for p in probs:
        c = pqdts_amd.Context.from_problem(p)
        start(c, p)
        ctxs.append(c)
        b.admit(c, 2, 300, p.eps, 1, 0, 0)
    res = b.run()
    for i, (p, c) in enumerate(zip(probs, ctxs)):
        Xa, conva, eva = alone[i]
        Xb = c.get("X")
        assert np.array_equal(Xa, Xb), f"problem {p.name}: max diff {np.max(np.abs(Xa - Xb)):.3e}"
        assert res[i]["conv"] == conva
        evb = b.events(i)
        assert [k for k, _ in evb] == [k for k, _ in eva]
        assert all(np.array_equal(x[1], y[1]) for x, y in zip(evb, eva))


def test_split_gram_batch_equals_alone_bitwise():
    """Problems whose Gram product is split over band chunks (partials summed by a
    second pass) beside unsplit ones: still bit-identical alone and batched."""
    probs = problems()
    splits = [3, 1, 2, 0]

    def setup_of(i):
        def setup(c):
            c.set_hvp_path("G")
            c.set_hvp_split(splits[i])
        return setup

    alone = [solve_alone(p, setup=setup_of(i)) for i, p in enumerate(probs)]
    b = pqdts_amd.Batch(8)
    ctxs = []
    for i, p in enumerate(probs):
        c = pqdts_amd.Context.from_problem(p)
        setup_of(i)(c)
        assert c.info()["hvp_split"] == (splits[i] or c.info()["hvp_split"])
        start(c, p)
        ctxs.append(c)
        b.admit(c, 2, 300, p.eps, 1, 0, 0)
    res = b.run()
    for i, (p, c) in enumerate(zip(probs, ctxs)):
        Xa, conva, _ = alone[i]
        assert np.array_equal(Xa, c.get("X")), f"problem {p.name}"
        assert res[i]["conv"] == conva


@pytest.mark.parametrize("precision", [0, 1])
def test_launch_forms_same_bits(precision):
    """FP64 band products of launches too small to fill the GPU use the latency
    form of the epilogue, large launches the occupancy form. A small problem
    batched beside a large one (every product launch then has over 1024
    workgroups) computes the same bits as alone, in FP64 and mixed precision."""
    small = workloads.random_problem(130, 60, 260, 0.6, 1e-2, 3e6, seed=13, eps=1e-7, name="small")
    big = workloads.random_problem(3200, 1400, 6400, 0.8, 1e-3, 1e7, seed=14, eps=1e-7, name="big")

    def run(probs):
        b = pqdts_amd.Batch(4)
        ctxs = []
        for p in probs:
            c = pqdts_amd.Context.from_problem(p)
            start(c, p)
            ctxs.append(c)
            b.admit(c, 2, 2, p.eps, 1, 0, precision)
        b.run()
        return [c.get("X") for c in ctxs]

    (xa,) = run([small])
    xb, _ = run([small, big])
    assert np.array_equal(xa, xb), f"max diff {np.max(np.abs(xa - xb)):.3e}"


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_same_bits(monkeypatch, precision, mode):
    """Launches replayed from captured hipGraphs give the bits and the events of the
    eager launches, over whole solves. PQDTS_GRAPHS=1 captures the CG rounds in
    blocks of 8; PQDTS_GRAPHS=2 captures a Newton iteration as two graphs around
    the host-driven line search. Graphs are replayed while the set of problems and
    their stages stay the same and captured anew when they change."""
    probs = problems()

    def run(graphs):
        monkeypatch.setenv("PQDTS_GRAPHS", graphs)
        b = pqdts_amd.Batch(8)
        ctxs = []
        for p in probs:
            c = pqdts_amd.Context.from_problem(p)
            start(c, p)
            ctxs.append(c)
            b.admit(c, 2, 300, p.eps, 1, 0, precision)
        res = b.run()
        ev = [list(b.events(i)) for i in range(len(probs))]
        return [c.get("X") for c in ctxs], [res[i] for i in range(len(probs))], ev

    xa, ra, ea = run("0")
    xb, rb, eb = run(mode)
    for p, a, b_ in zip(probs, xa, xb):
        assert np.array_equal(a, b_), f"problem {p.name}"
    assert ra == rb
    assert ea == eb


def test_batch_membership_rules():
    """A context sits in at most one batch slot; while it does, the Hessian-path
    settings (which rebuild its device descriptor) are refused; release frees it;
    destroying a context before its batch leaves the batch usable."""
    p = problems()[1]
    c = pqdts_amd.Context.from_problem(p)
    start(c, p)
    b1, b2 = pqdts_amd.Batch(2), pqdts_amd.Batch(2)
    s1 = b1.admit(c, 2, 5, p.eps, 1, 0, 0)
    with pytest.raises(pqdts_amd.PqdtsError):
        b2.admit(c, 2, 5, p.eps, 1, 0, 0)
    with pytest.raises(pqdts_amd.PqdtsError):
        c.set_hvp_split(2)
    with pytest.raises(pqdts_amd.PqdtsError):
        c.set_hvp_path("F")
    b1.release(s1)
    c.set_hvp_path("G")
    s2 = b2.admit(c, 2, 5, p.eps, 1, 0, 0)
    b2.run()
    b2.release(s2)
    # destroyed while still admitted: the batch drops the slot, then closes cleanly
    d = pqdts_amd.Context.from_problem(p)
    start(d, p)
    b1.admit(d, 2, 5, p.eps, 1, 0, 0)
    d.close()
    assert pqdts_amd._lib.load().pqdts_batch_active(b1.h) == 0
    b1.close()
    b2.close()


def test_continuous_admission_matches_alone():
    probs = problems()
    alone = [solve_alone(p) for p in probs]
    b = pqdts_amd.Batch(3)
    ctxs = {}
    pending = list(range(len(probs)))
    done = {}
    steps = 0
    while pending or b.active():
        # admit while slots are free
        while pending and b.active() < 3:
            i = pending.pop(0)
            c = pqdts_amd.Context.from_problem(probs[i])
            start(c, probs[i])
            s = b.admit(c, 2, 300, probs[i].eps, 1, 0, 0)
            ctxs[s] = (i, c)
        for s in b.step():
            i, c = ctxs.pop(s)
            done[i] = c.get("X")
            b.release(s)
        steps += 1
        assert steps < 5000
    for i in range(len(probs)):
        assert np.array_equal(done[i], alone[i][0]), probs[i].name


@pytest.mark.parametrize("path", ["F", "G"])
def test_mixed_precision_within_1e6_of_fp64(path):
    p = workloads.config1()
    eps = 1e-11
    out = {}
    for prec in (0, 1):
        c = pqdts_amd.Context.from_problem(p)
        c.set_hvp_path(path)
        start(c, p)
        b = pqdts_amd.Batch(1)
        b.admit(c, 2, 3000, eps, 1, 0, prec)
        r = b.run()[0]
        its = [v for k, v in b.events(0) if k == 0]
        out[prec] = (c.get("X"), its[-1][2], r)
    (X64, L64, r64), (X32, L32, r32) = out[0], out[1]
    dpi = float(np.max(np.abs(X64 - X32)))
    gap = abs(L32 - L64) / L64
    print(f"path {path}: ||dPi||inf = {dpi:.3e}, rel objective gap = {gap:.3e}, "
          f"newton {r64['newton']} vs {r32['newton']}, pcg {r64['pcg']} vs {r32['pcg']}")
    assert r64["conv"] < eps and r32["conv"] < eps
    assert dpi <= 1e-6
    assert gap <= 1e-9


def test_cu_restricted_batch_matches_alone():
    """A batch limited to a subset of the compute units gives the same bits."""
    p = problems()[1]
    Xa, conva, _ = solve_alone(p)
    c = pqdts_amd.Context.from_problem(p)
    start(c, p)
    b = pqdts_amd.Batch(2)
    b.set_cus(range(8, 40))
    b.admit(c, 2, 300, p.eps, 1, 0, 0)
    r = b.run()[0]
    assert np.array_equal(c.get("X"), Xa)
    assert r["conv"] == conva
    with pytest.raises(pqdts_amd.PqdtsError):
        b.set_cus(range(4))  # only an empty batch can be restricted


# Problems of the benchmark mix (workloads.batch_mix(256), by index) whose optimum FP64
# determines at eps = 1e-11: two FP64 solves that differ only in the Hessian product's
# summation order agree there to <= 6e-7 (asserted below). Small gamma or large M leave
# Pi undetermined far beyond 1e-6 at any eps FP64 can reach, for FP64 itself as much as
# for mixed precision, so the 1e-6 bar is asserted where it is meaningful.
DETERMINED_MIX = [4, 10, 13, 34, 35, 45, 49, 94, 102, 139, 206, 225]


def test_mixed_precision_meets_1e6_on_mix_problems():
    """Batched mixed precision (FP32 Hessian products inside CG, FP64 everywhere else)
    against batched FP64 (bit-identical to the plain solver alone) at eps = 1e-11 on 12
    problems of the uneven mix: every problem converged in both, ||dPi||inf <= 1e-6 and
    relative objective gap <= 1e-9; the FP64-vs-FP64 spread is within the same bar."""
    allp = workloads.batch_mix(256, eps=1e-11)
    probs = [allp[i] for i in DETERMINED_MIX]
    eps = 1e-11

    def run(prec, path="auto"):
        b = pqdts_amd.Batch(len(probs))
        ctxs = []
        for p in probs:
            c = pqdts_amd.Context.from_problem(p)
            c.set_hvp_path(path)
            start(c, p)
            ctxs.append(c)
            b.admit(c, 2, 5000, eps, 1, 0, prec, record=False)
        res = b.run()
        return [c.get("X") for c in ctxs], [c.L("X") for c in ctxs], [res[i] for i in range(len(probs))]

    X64, L64, r64 = run(0)
    Xm, Lm, rm = run(1)
    XF, _, rF = run(0, "F")
    for k, p in enumerate(probs):
        assert r64[k]["conv"] < eps and rm[k]["conv"] < eps and rF[k]["conv"] < eps, p.name
        dpi = float(np.max(np.abs(X64[k] - Xm[k])))
        floor = float(np.max(np.abs(X64[k] - XF[k])))
        gap = abs(Lm[k] - L64[k]) / abs(L64[k])
        print(f"{p.name} M={p.M} N={p.N} gamma={p.gamma:.1e}: mixed ||dPi||inf {dpi:.2e}, "
              f"fp64-vs-fp64 {floor:.2e}, gap {gap:.1e}")
        assert dpi <= 1e-6 and floor <= 1e-6 and gap <= 1e-9, p.name
