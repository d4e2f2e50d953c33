"""GPU parity at the benchmark configurations' real shapes, through the C ABI.

* mid-size (M=2000, N=1000, D=4000, the plain FP64 case): every operator against
  the CPU oracle in both spaces, for the two Hessian-product paths (F products;
  Gram matrix split two ways over band chunks, the shape's automatic choice).
* large (M=20000, N=10000, D=40000): the products and the stencil are separable
  by column, so a 128-column slice of Pi on the same F is compared with the
  oracle, with P noise-free (P = F Pi_true) so that L(Pi_true) = gamma *
  smoothing(Pi_true). At the full size, properties that fix the operators
  without an oracle: rows of doproj on the simplex, <u, Hv> = <Hu, v>,
  dL(X + v) - dL(X) = H v (the objective is quadratic), and with P = 0,
  L(X) = <X, H X> / 2.
* mix (256 uneven problems, every 8th): batched FP64 and batched mixed precision
  are bit-identical to each problem solved alone, and the alone solves of the
  small ones follow the oracle's first Newton iterations.

Bar as elsewhere: relative 1e-12 of the largest term for arrays, 1e-12 for
scalars (the GPU sums in blocked order, the oracle sequentially).
"""
import numpy as np
import pytest
import scipy.sparse as sps
from scipy.stats import binom

import oracle
import pqdts_amd
from pqdts_amd import workloads

pytestmark = pytest.mark.gpu


def close_arr(a, b, rtol=1e-12):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    err = float(np.max(np.abs(a - b)))
    assert err <= rtol * scale, f"max err {err:.3e} vs scale {scale:.3e}"


def close(a, b, rtol=1e-12):
    assert abs(a - b) <= rtol * max(abs(b), 1e-300), (a, b)


def simplex_rows(rng, M, N):
    X = rng.random((M, N)) ** 3
    return X / X.sum(1, keepdims=True)


# --------------------------------------------------------------- mid-size case
@pytest.fixture(scope="module")
def cfg2():
    return workloads.config2()


@pytest.mark.parametrize("path", ["F", "G2"])
def test_cfg2_operators_match_oracle(cfg2, path):
    p = cfg2
    ctx = pqdts_amd.Context.from_problem(p)
    ctx.set_hvp_path(path[0])
    if path == "G2":
        ctx.set_hvp_split(2)
        assert ctx.info()["hvp_split"] == 2
    orc = oracle.Oracle.from_problem(p)
    rng = np.random.default_rng(21)

    def both(slot, a):
        ctx.put(slot, a)
        orc.buf(slot)[:] = a

    X = simplex_rows(rng, p.M, p.N)
    V = rng.normal(size=(p.M, p.N))
    for full in (True, False):
        both("X", X)
        if not full:
            ctx.fullspace = False
            orc.set_fullspace(False)
            ctx.transform()
            orc.transform()
            assert np.array_equal(ctx.maxr, orc.maxr)
        close(ctx.L("X"), orc.L("X"))
        ctx.dL("X", "G")
        orc.dL("X", "G")
        close_arr(ctx.get("G"), orc.buf("G"))
        close_arr(ctx.get_O(), orc.O)
        both("XTMP", V)
        ctx.BF("XTMP", "HP")
        orc.BF("XTMP", "HP")
        close_arr(ctx.get("HP"), orc.buf("HP"))
        orc.Bdiag()
        close_arr(ctx.Bdiag(), orc.diag)
    # doproj2 and conv_test in the reduced space, then doproj (full space)
    Y = orc.buf("X").copy() + rng.normal(size=(p.M, p.N)) * 1e-3
    both("X", Y)
    close(ctx.doproj2("X"), orc.doproj2("X"))
    close_arr(ctx.get("X"), orc.buf("X"), 1e-13)
    close(ctx.conv_test("X", "G"), orc.conv_test("X", "G"), 1e-10)
    Y = rng.normal(size=(p.M, p.N)) * 0.01 + 1.0 / p.N
    both("X", Y)
    ctx.doproj("X")
    orc.doproj("X")
    close_arr(ctx.get("X"), orc.buf("X"), 1e-13)


# ------------------------------------------------------------------ large case
LM, LN, LD = 20000, 10000, 40000


@pytest.fixture(scope="module")
def cfg3_F():
    return workloads.poisson_rows(np.linspace(0.0, 0.9 * LM, LD), LM)


def test_cfg3_column_slice_matches_oracle(cfg3_F):
    """The large case's F (M=20000, D=40000) with 128 columns of Pi: L, dL, BF and
    Bdiag against the oracle; L(Pi_true) = gamma * smoothing(Pi_true) for P = F Pi_true."""
    F, ns, gamma = cfg3_F, 128, 1e-2
    # Pi_true's first 128 columns (eta = 0.9 loss, no saturation below N-1);
    # rows above 400 photons put no weight on outcomes below 128 at this eta
    i = np.arange(400)[:, None]
    pit = np.zeros((LM, ns))
    pit[:400] = binom.pmf(np.arange(ns)[None, :], i, 0.9)
    P = np.asarray(F[:, :400] @ pit[:400])
    Pc = sps.coo_matrix(P)
    Fc = F.tocoo()
    ctx = pqdts_amd.Context(LM, ns, LD, (Fc.row, Fc.col, Fc.data), (Pc.row, Pc.col, Pc.data), gamma)
    info = ctx.info()
    assert info["gram"]
    orc = oracle.Oracle(LM, ns, LD, (Fc.row, Fc.col, Fc.data), (Pc.row, Pc.col, Pc.data), gamma)
    ctx.put("X", pit)
    smooth = gamma * float(np.sum((pit[:-1] - pit[1:]) ** 2))
    lg = ctx.L("X")
    assert abs(lg - smooth) <= 1e-10 * smooth, (lg, smooth)
    rng = np.random.default_rng(22)
    X = simplex_rows(rng, LM, ns)
    ctx.put("X", X)
    orc.buf("X")[:] = X
    close(ctx.L("X"), orc.L("X"))
    ctx.dL("X", "G")
    orc.dL("X", "G")
    close_arr(ctx.get("G"), orc.buf("G"))
    V = rng.normal(size=(LM, ns))
    ctx.put("XTMP", V)
    orc.buf("XTMP")[:] = V
    ctx.BF("XTMP", "HP")
    orc.BF("XTMP", "HP")
    close_arr(ctx.get("HP"), orc.buf("HP"))
    orc.Bdiag()
    close_arr(ctx.Bdiag(), orc.diag)


def test_cfg3_full_size_invariants(cfg3_F):
    """M=20000, N=10000, D=40000 (Pi and every work array 1.6 GB, ~18 GB on the
    device), P = 0: properties of the operators at the full shape."""
    Fc = cfg3_F.tocoo()
    gamma = 1e-2
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    ctx = pqdts_amd.Context(LM, LN, LD, (Fc.row, Fc.col, Fc.data), empty, gamma)
    rng = np.random.default_rng(23)
    # doproj: rows on the simplex
    ctx.put("X", rng.normal(size=(LM, LN)) * 1e-3 + 1.0 / LN)
    ctx.doproj("X")
    X = ctx.get("X")
    assert X.min() >= 0.0 and np.max(np.abs(X.sum(1) - 1.0)) <= 1e-12
    # L(X) = <X, H X> / 2 when P = 0
    ctx.pcopy("X", "XTMP")
    ctx.BF("XTMP", "HP")
    close(ctx.L("X"), 0.5 * ctx.pdot("X", "HP"))
    # symmetry: <H u, v> = <u, H v>
    U = rng.normal(size=(LM, LN))
    ctx.put("SN", U)
    ctx.put("Z", rng.normal(size=(LM, LN)))
    ctx.pcopy("SN", "XTMP")
    ctx.BF("XTMP", "HP")
    huv = ctx.pdot("HP", "Z")
    ctx.pcopy("Z", "XTMP")
    ctx.BF("XTMP", "HP")          # HP = H v, v = Z
    uhv = ctx.pdot("SN", "HP")
    # rounding of a dot product is relative to |Hu| |v|, not to the dot itself
    scale = np.sqrt(ctx.pdot("HP", "HP") * ctx.pdot("SN", "SN"))
    assert abs(huv - uhv) <= 1e-13 * scale, (huv, uhv, scale)
    # dL(X + v) - dL(X) = H v, v = Z: dL through the F products, H v through G
    ctx.dL("X", "G")
    ctx.pcopy("G", "SN")          # SN = dL(X)
    ctx.paxpy(1.0, "Z", "X")      # X = X + v
    ctx.dL("X", "G")
    ctx.paxpy(-1.0, "SN", "G")    # G = dL(X + v) - dL(X)
    ctx.paxpy(-1.0, "HP", "G")    # minus H v
    res, ref = ctx.pdot("G", "G"), ctx.pdot("HP", "HP")
    assert np.sqrt(res) <= 1e-11 * np.sqrt(ref), (np.sqrt(res), np.sqrt(ref))


# --------------------------------------------------------------------- the mix
@pytest.fixture(scope="module")
def mix32():
    return workloads.batch_mix(256, eps=1e-7)[::8]


def _alone(p, maxiter):
    c = pqdts_amd.Context.from_problem(p)
    c.pset(1.0 / p.N, "X")
    conv, ev = c.solve(2, maxiter, p.eps, 1, 0)
    return c.get("X"), conv, ev


@pytest.mark.parametrize("precision", [0, 1])
def test_mix_batched_equals_alone_bitwise(mix32, precision):
    """32 uneven problems of the mix in one batch, capped at 12 Newton iterations
    per stage: every problem's Pi, conv and event stream are the same bits as
    when it runs alone (FP64: the plain pqdts_solve; mixed: a batch of one)."""
    maxiter = 12
    if precision == 0:
        alone = [_alone(p, maxiter) for p in mix32]
    else:
        alone = []
        for p in mix32:
            c = pqdts_amd.Context.from_problem(p)
            c.pset(1.0 / p.N, "X")
            b = pqdts_amd.Batch(1)
            b.admit(c, 2, maxiter, p.eps, 1, 0, 1)
            r = b.run()[0]
            alone.append((c.get("X"), r["conv"], b.events(0)))
            b.close()
    b = pqdts_amd.Batch(len(mix32))
    ctxs = []
    for p in mix32:
        c = pqdts_amd.Context.from_problem(p)
        c.pset(1.0 / p.N, "X")
        ctxs.append(c)
        b.admit(c, 2, maxiter, p.eps, 1, 0, precision)
    res = b.run()
    for i, (p, c) in enumerate(zip(mix32, ctxs)):
        Xa, conva, eva = alone[i]
        assert np.array_equal(c.get("X"), Xa), f"{p.name} M={p.M} N={p.N}"
        assert res[i]["conv"] == conva
        evb = b.events(i)
        assert [k for k, _ in evb] == [k for k, _ in eva]
        assert all(np.array_equal(x[1], y[1]) for x, y in zip(evb, eva))


def test_mix_small_problems_follow_oracle(mix32):
    """The plain FP64 solve of the mix's problems with M <= 300 against the oracle
    over the first two Newton iterations: same line-search alphas and trial
    counts, objective and conv to 1e-10."""
    small = [p for p in mix32 if p.M <= 300]
    assert len(small) >= 6
    for p in small:
        _, _, ev = _alone(p, 2)
        o = oracle.Oracle.from_problem(p)
        o.buf("X")[:] = 1.0 / p.N
        _, evo = o.solve(2, 2, p.eps, 1, 0)
        a = [v for k, v in ev if k == 0]
        b = [v for k, v in evo if k == 0]
        assert len(a) == len(b), p.name
        for x, y in zip(a, b):
            assert x[7] == y[7], (p.name, x, y)
            close(x[2], y[2], 1e-10)
            close(x[4], y[4], 1e-8)
        la = [v for k, v in ev if k == 2]
        lb = [v for k, v in evo if k == 2]
        assert len(la) == len(lb), p.name
