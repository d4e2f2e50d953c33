"""GPU parity of every operator against the CPU oracle, through the C ABI.

Bar: the operators are FP64 with different (blocked / tree) summation orders
than the oracle's sequential loops, so values agree to a few ulps of the
largest term: relative 1e-12 for scalars, 1e-12 of the array's max-abs for
arrays; discrete outputs (maxr, active set) are identical."""
import numpy as np
import pytest
import scipy.sparse as sps

import oracle
import pqdts_amd
from pqdts_amd import workloads

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def close_arr(a, b, rtol=RTOL):
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b))
    assert err <= rtol * scale, f"max err {err:.3e} vs scale {scale:.3e}"


def close(a, b, rtol=RTOL):
    assert abs(a - b) <= rtol * max(abs(b), 1e-300), (a, b)


def ragged_problem(M=77, N=131, D=150, gamma=2e-3, seed=0):
    """Banded F with uneven spans, one empty row, duplicates in the COO."""
    rng = np.random.default_rng(seed)
    alpha2 = np.linspace(0, 0.8 * M, D)
    F = workloads.poisson_rows(alpha2, M).tocoo()
    rows, cols, vals = list(F.row), list(F.col), list(F.data)
    keep = [i for i in range(len(rows)) if rows[i] != 5]  # empty row 5
    rows, cols, vals = [rows[i] for i in keep], [cols[i] for i in keep], [vals[i] for i in keep]
    rows.append(rows[10]); cols.append(cols[10]); vals.append(0.123)  # duplicate: last wins
    P = rng.random((D, N))
    P /= P.sum(1, keepdims=True)
    Pc = sps.coo_matrix(P)
    coo = (np.array(rows), np.array(cols), np.array(vals))
    ctx = pqdts_amd.Context(M, N, D, coo, (Pc.row, Pc.col, Pc.data), gamma)
    orc = oracle.Oracle(M, N, D, coo, (Pc.row, Pc.col, Pc.data), gamma)
    return ctx, orc, M, N


@pytest.fixture(params=["cfg1", "ragged"])
def pair(request):
    if request.param == "cfg1":
        p = workloads.config1()
        return pqdts_amd.Context.from_problem(p), oracle.Oracle.from_problem(p), p.M, p.N
    return ragged_problem()


def set_both(ctx, orc, slot, arr):
    ctx.put(slot, arr)
    orc.buf(slot)[:] = arr


def simplex_rows(rng, M, N):
    X = rng.random((M, N)) ** 3
    return X / X.sum(1, keepdims=True)


def test_gpu_visible():
    assert pqdts_amd.device_count() >= 1


@pytest.mark.parametrize("path", ["F", "G", "G3"])
@pytest.mark.parametrize("full", [True, False])
def test_L_dL_BF(pair, full, path):
    """path F: BF as 2 F^T (F p); path G: BF with the banded Gram matrix F^T F;
    G3: the Gram product split three ways over band chunks."""
    ctx, orc, M, N = pair
    ctx.set_hvp_path(path[0])
    if path == "G3":
        ctx.set_hvp_split(3)
        assert ctx.info()["hvp_split"] == 3
    rng = np.random.default_rng(1)
    set_both(ctx, orc, "X", simplex_rows(rng, M, N))
    if not full:
        ctx.fullspace = False
        orc.set_fullspace(False)
        ctx.transform()
        orc.transform()
        assert np.array_equal(ctx.maxr, orc.maxr)
        close_arr(ctx.get("X"), orc.buf("X"), 0)
    close(ctx.L("X"), orc.L("X"))
    # the implicit entry is rebuilt as 1 - sum(others): a few ulps of 1.0
    assert np.max(np.abs(ctx.get("X") - orc.buf("X"))) <= 1e-14
    ctx.dL("X", "G")
    orc.dL("X", "G")
    close_arr(ctx.get("G"), orc.buf("G"))
    # O = F X as left by the gradient
    close_arr(ctx.get_O(), orc.O)
    V = rng.normal(size=(M, N))
    set_both(ctx, orc, "XTMP", V)
    ctx.BF("XTMP", "HP")
    orc.BF("XTMP", "HP")
    close_arr(ctx.get("HP"), orc.buf("HP"))
    assert ctx.info()["gram"] == (path[0] == "G")
    orc.Bdiag()
    close_arr(ctx.Bdiag(), orc.diag)


def test_pdot_and_blas1(pair):
    ctx, orc, M, N = pair
    rng = np.random.default_rng(2)
    A, B = rng.normal(size=(M, N)), rng.normal(size=(M, N))
    set_both(ctx, orc, "X", A)
    set_both(ctx, orc, "G", B)
    close(ctx.pdot("X", "G"), orc.pdot("X", "G"))
    ctx.pgaxpy2(2.0, "X", 0.5, "G", "Z")
    close_arr(ctx.get("Z"), 0.5 * B + 2.0 * A, 1e-15)
    ctx.paxpy(-0.25, "X", "G")
    close_arr(ctx.get("G"), B - 0.25 * A, 1e-15)
    ctx.pcopy_scale(-1.0, "X", "SN")
    assert np.array_equal(ctx.get("SN"), -A)
    ctx.pset(3.0, "HP")
    assert np.all(ctx.get("HP") == 3.0)


def test_doproj_condat(pair):
    ctx, orc, M, N = pair
    rng = np.random.default_rng(3)
    Y = rng.normal(size=(M, N)) * 0.2 + 1.0 / N
    Y[0] = 1.0 / N  # already on the simplex
    Y[1] = -1.0     # all equal and negative
    set_both(ctx, orc, "X", Y)
    ctx.doproj("X")
    orc.doproj("X")
    close_arr(ctx.get("X"), orc.buf("X"), 1e-13)
    assert np.allclose(ctx.get("X").sum(1), 1.0, atol=1e-13)


def test_simplexproj_rows_vs_reference_vectors():
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "condat_vectors.npz"))
    n = len([k for k in z.files if k.startswith("y")])
    for i in range(n):
        y, x = z[f"y{i}"], z[f"x{i}"]
        got = pqdts_amd.simplexproj_rows(y[None, :])[0]
        assert np.max(np.abs(got - x)) <= 1e-13 * max(1.0, np.max(np.abs(y))), i


def test_simplexproj_condat_symbol_bit_exact_vs_reference_vectors():
    """The reference's own foreign function, simplexproj_condat_(y, x, &n)
    (condat_simplexproj.c:60), through ctypes with its by-reference signature:
    x bit-identical to the reference build's output on every golden vector
    (random, ties, all-negative, n = 1, n = 4097), y left unchanged."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "condat_vectors.npz"))
    n = len([k for k in z.files if k.startswith("y")])
    assert n >= 30
    for i in range(n):
        y, x = z[f"y{i}"].copy(), z[f"x{i}"]
        y0 = y.copy()
        got = pqdts_amd.simplexproj_condat(y)
        assert np.array_equal(got, x), (i, float(np.max(np.abs(got - x))))
        assert np.array_equal(y, y0)


def test_doproj2_conv_transform(pair):
    ctx, orc, M, N = pair
    rng = np.random.default_rng(4)
    X = simplex_rows(rng, M, N)
    set_both(ctx, orc, "X", X)
    ctx.fullspace = False
    orc.set_fullspace(False)
    ctx.transform()
    orc.transform()
    Y = orc.buf("X").copy() + rng.normal(size=(M, N)) * 0.01
    set_both(ctx, orc, "X", Y)
    close(ctx.doproj2("X"), orc.doproj2("X"), 1e-12)
    close_arr(ctx.get("X"), orc.buf("X"), 1e-13)
    close(ctx.conv_test("X", "G"), orc.conv_test("X", "G"), 1e-10)
    assert ctx.fullspace and orc.fullspace


@pytest.mark.parametrize("stage", [1, 2])
def test_line_search(pair, stage):
    ctx, orc, M, N = pair
    rng = np.random.default_rng(5)
    set_both(ctx, orc, "X", simplex_rows(rng, M, N))
    if stage == 2:
        for o in (ctx, orc):
            if o is ctx:
                o.fullspace = False
            else:
                o.set_fullspace(False)
            o.transform()
    S = rng.normal(size=(M, N)) * 0.05
    if stage == 2:
        S[np.arange(M), orc.maxr] = 0.0
    set_both(ctx, orc, "SN", S)
    if stage == 1:
        a, dl = ctx.lalpha_min_auglag()
        ao, dlo = orc.lalpha_min_auglag()
    else:
        a, dl = ctx.lalpha_min2()
        ao, dlo = orc.lalpha_min2()
    assert a == ao
    close(dl, dlo, 1e-6)  # finite-difference slope: (L(dx) - L(0)) / 1e-5 amplifies rounding


def test_bad_arguments_raise():
    p = workloads.config1()
    ctx = pqdts_amd.Context.from_problem(p)
    with pytest.raises(pqdts_amd.PqdtsError):
        ctx.L(17)
    ctx.fullspace = False  # no maxr yet: reduced-space ops must refuse, not fault
    with pytest.raises(pqdts_amd.PqdtsError, match="maxr"):
        ctx.dL("X", "G")
    with pytest.raises(pqdts_amd.PqdtsError):
        pqdts_amd.Context(3, 2, 2, (np.array([0]), np.array([5]), np.array([1.0])),
                          (np.array([0]), np.array([0]), np.array([1.0])), 0.0)
