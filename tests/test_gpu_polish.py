"""Polish: the primal-dual active-set Newton continuation after conv < eps.

The reference stops at conv < eps (pqdts.f90:1034-1067, 2134). conv weights each
KKT residual by X, so on the uneven mix's small-gamma problems points 1e-3 apart
pass eps = 1e-7, and two solves that differ only in rounding end that far apart.
The polish continues to the problem's minimiser, which is unique (the problem is
strictly convex). So its result is checked against the KKT conditions
(oracle/qp.py), against the oracle's exact solution, and across precisions,
batching and summation orders.
"""
import numpy as np
import pytest

import pqdts_amd
from oracle import qp
from pqdts_amd import workloads

pytestmark = pytest.mark.gpu

# problems of the 256-problem mix (M, gamma): 193 (71, 1.3e-5), 162 (82, 7.7e-3),
# 245 (88, 7.6e-3), 229 (78, 1.9e-4), 214 (78, 5.3e-4), 170 (229, 3.5e-4)
MIX = [193, 162, 245, 229, 214, 170]
EPS = 1e-7


@pytest.fixture(scope="module")
def probs():
    return workloads.batch_mix(256, eps=EPS, only=MIX)


def run_batch(probs, precision, path="auto"):
    b = pqdts_amd.Batch(len(probs))
    ctxs = []
    for p in probs:
        c = pqdts_amd.Context.from_problem(p)
        c.set_hvp_path(path)
        c.set_polish()
        c.put("X", np.full((p.M, p.N), 1.0 / p.N))
        ctxs.append(c)
        b.admit(c, 2, 5000, EPS, 1, 0, precision, record=False)
    res = b.run()
    for slot in list(res):
        b.release(slot)
    b.close()
    out = [(c.get("X"), res[k], c.polish_result(), c.L("X")) for k, c in enumerate(ctxs)]
    for c in ctxs:
        c.close()
    return out


def test_polish_reaches_the_certified_minimiser_in_both_precisions(probs):
    """FP64 and mixed precision (FP32 Hessian products in the minimisation), both
    polished: the oracle's exact dense active-set method, started from the FP64
    result, ends at a point that passes the KKT certificate (the minimiser); both
    results are within 1e-8 of it, and the precisions agree with an objective gap
    <= 1e-12: the 1e-6 bar of the precision form, with margin."""
    r64 = run_batch(probs, 0)
    rm = run_batch(probs, 1)
    for p, (X64, res64, pol64, L64), (Xm, resm, polm, Lm) in zip(probs, r64, rm):
        tag = f"{p.name} M={p.M} N={p.N} gamma={p.gamma:.1e}"
        print(tag, "fp64", pol64, "mixed", polm)
        assert res64["conv"] < EPS and resm["conv"] < EPS, tag
        assert pol64["status"] == 1 and polm["status"] == 1, tag
        Xe, ok = qp.exact(p.F, p.P, p.gamma, X0=X64)
        assert ok and qp.is_minimiser(p.F, p.P, p.gamma, Xe), tag
        for X in (X64, Xm):
            k = qp.kkt(p.F, p.P, p.gamma, X)
            assert k["x_min"] >= 0.0 and k["rowsum_err"] <= 1e-12, (tag, k)
            assert np.max(np.abs(X - Xe)) <= 1e-8, (tag, float(np.max(np.abs(X - Xe))))
        assert abs(Lm - L64) <= 1e-12 * abs(L64), tag


def test_polish_batched_equals_alone_bitwise(probs):
    """An FP64 polish computes the same bits alone (pqdts_solve) and in a batch,
    with the same iteration and CG-round counts."""
    batched = run_batch(probs[:3], 0)
    for p, (Xb, _, polb, _) in zip(probs[:3], batched):
        c = pqdts_amd.Context.from_problem(p)
        c.set_polish()
        c.put("X", np.full((p.M, p.N), 1.0 / p.N))
        c.solve(2, 5000, EPS, 1, 0)
        assert np.array_equal(c.get("X"), Xb), p.name
        assert c.polish_result() == polb, p.name
        c.close()


def test_polish_removes_the_summation_order_spread(probs):
    """Two FP64 solves whose Hessian products sum in different orders (Gram matrix
    vs the two F products) end ~1e-3 apart at conv < 1e-7 on small-gamma problems;
    polished they agree to 1e-8."""
    idx = [0, 3]  # gamma 1.3e-5 and 1.9e-4
    sub = [probs[i] for i in idx]
    rg = run_batch(sub, 0, "G")
    rf = run_batch(sub, 0, "F")
    for p, a, b in zip(sub, rg, rf):
        assert a[2]["status"] == 1 and b[2]["status"] == 1, p.name
        assert np.max(np.abs(a[0] - b[0])) <= 1e-8, p.name


def test_polish_at_a_vertex_minimiser():
    """A problem whose minimiser is a vertex of every row's simplex (noise-free
    P = F Pi_true with one-hot rows, gamma = 0, F of full column rank): the last
    faces have no free entries (r = p = 0, p.Hp = 0). The polish ends at Pi_true
    with no 0/0 in its CG step, in both precisions."""
    import scipy.sparse as sps
    rng = np.random.default_rng(8)
    M, N, D = 12, 5, 30
    F = sps.csr_matrix(rng.random((D, M)) + 0.1)
    pi = np.zeros((M, N))
    pi[np.arange(M), rng.integers(0, N, M)] = 1.0
    p = workloads.Problem("vertex", M, N, D, F, np.ascontiguousarray(F @ pi), 0.0, 1e-9, pi)
    for prec in (0, 1):
        (X, res, pol, _), = run_batch([p], prec)
        assert pol["status"] == 1 and np.all(np.isfinite(X)), (prec, pol)
        assert np.max(np.abs(X - pi)) <= 1e-9, prec


def test_polish_stopped_by_its_budget_returns_the_minimisation_result(probs):
    """A polish that reaches its CG budget (here 0.05 x the minimisation's own PCG
    rounds) reports status 2 and leaves X and conv exactly as the solve without a
    polish ends them, in both precisions."""
    p = probs[5]  # M = 229: its polish needs more rounds than the budget allows
    for prec in (0, 1):
        outs = []
        for budget in (None, 0.05):
            c = pqdts_amd.Context.from_problem(p)
            if budget is not None:
                c.set_polish(60, budget=budget)
            c.put("X", np.full((p.M, p.N), 1.0 / p.N))
            b = pqdts_amd.Batch(1)
            b.admit(c, 2, 5000, EPS, 1, 0, prec, record=False)
            res = b.run()[0]
            b.release(0)
            b.close()
            outs.append((c.get("X"), res["conv"], c.polish_result()))
            c.close()
        (X0, conv0, pol0), (X1, conv1, pol1) = outs
        assert pol0["status"] == 0 and pol1["status"] == 2, (prec, pol1)
        assert np.array_equal(X0, X1) and conv0 == conv1, prec
