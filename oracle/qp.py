"""The problem's exact minimiser and a KKT certificate -- TEST INFRASTRUCTURE ONLY.

The objective of pqdts (README.md:4-16; L at pqdts.f90:917-973, dL at 975-1032)

    f(Pi) = ||P - F Pi||^2 + gamma * sum_n sum_i (Pi[i,n] - Pi[i+1,n])^2,
    every row of Pi (M x N) on the probability simplex,

is a quadratic whose Hessian acts on each column n of Pi with the same M x M
matrix K = 2 F^T F + 2 gamma Lap (Lap: the path-graph Laplacian of the
second-difference stencil, pqdts.f90:1009-1019). With K positive definite the
problem is strictly convex, so a point satisfying its KKT conditions is its
unique minimiser:

    Pi >= 0,  Pi 1 = 1,  g = K Pi - 2 F^T P,
    mu[i,n] = g[i,n] + lambda_i >= 0,  mu[i,n] Pi[i,n] = 0.

Only tests/ use this module, as the checker of the library's polished
results; the product path never does. `exact` and `brute_force` are the
checkers; `face_polish` restates the library's direct polish iteration and is
itself checked against them (tests/test_oracle.py).
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sps


def hessian_block(F, gamma: float):
    """K = 2 F^T F + 2 gamma Lap, sparse (M x M)."""
    F = sps.csr_matrix(F)
    M = F.shape[1]
    K = (2.0 * (F.T @ F)).tocsr()
    if gamma != 0.0 and M > 1:
        d = np.zeros(M)
        d[:-1] += 1.0
        d[1:] += 1.0
        lap = sps.diags([d, -np.ones(M - 1), -np.ones(M - 1)], [0, 1, -1], format="csr")
        K = (K + 2.0 * gamma * lap).tocsr()
    return K


def gradient(F, P, gamma: float, X):
    return hessian_block(F, gamma) @ X - 2.0 * (sps.csr_matrix(F).T @ P)


def objective(F, P, gamma: float, X) -> float:
    R = P - sps.csr_matrix(F) @ X
    return float((R * R).sum() + gamma * ((X[:-1] - X[1:]) ** 2).sum())


def kkt(F, P, gamma: float, X, zero: float = 0.0) -> dict:
    """KKT residuals of X. Entries > zero are free; lambda_i is the mean of -g over
    the free entries of row i (the multiplier that makes their mu vanish on
    average). Returns the most negative entry, the largest row-sum error, the
    largest |mu| on free entries (stationarity), the most negative mu on the zero
    entries (dual feasibility), all relative to the gradient's scale."""
    X = np.asarray(X, dtype=np.float64)
    g = gradient(F, P, gamma, X)
    scale = float(np.abs(g).max()) or 1.0
    free = X > zero
    lam = np.empty(X.shape[0])
    for i in range(X.shape[0]):
        lam[i] = -g[i][free[i]].mean() if free[i].any() else -g[i].min()
    mu = g + lam[:, None]
    return {"x_min": float(X.min()), "rowsum_err": float(np.abs(X.sum(1) - 1.0).max()),
            "stationarity": float(np.abs(mu[free]).max()) / scale if free.any() else 0.0,
            "dual_min": float(mu[~free].min()) / scale if (~free).any() else 0.0,
            "free": int(free.sum())}


def is_minimiser(F, P, gamma, X, tol: float = 1e-9) -> bool:
    r = kkt(F, P, gamma, X)
    return r["x_min"] >= 0.0 and r["rowsum_err"] <= 1e-12 and r["stationarity"] <= tol and r["dual_min"] >= -tol


def _face_factors(K, free):
    """Per column n, (I_n, K_II^-1) for the free rows I_n, and the Schur complement
    S = sum_n E_n K_II^-1 E_n^T of the row-sum constraints."""
    M, N = free.shape
    S = np.zeros((M, M))
    inv = []
    for n in range(N):
        I = np.nonzero(free[:, n])[0]
        if I.size == 0:
            inv.append((I, None))
            continue
        Ki = np.linalg.inv(K[np.ix_(I, I)])
        S[np.ix_(I, I)] += Ki
        inv.append((I, Ki))
    return inv, S


def _face_solve(inv, S, B, c):
    """x[I_n, n] = K_II^-1 (B[I_n, n] - lambda[I_n]) with the row sums of x equal to c."""
    M, N = B.shape
    rhs = -np.asarray(c, dtype=np.float64).copy()
    for n, (I, Ki) in enumerate(inv):
        if I.size:
            rhs[I] += Ki @ B[I, n]
    lam = np.linalg.solve(S, rhs)
    X = np.zeros((M, N))
    for n, (I, Ki) in enumerate(inv):
        if I.size:
            X[I, n] = Ki @ (B[I, n] - lam[I])
    return X, lam


def _face_solution(K, B, free):
    """Minimiser of the quadratic on the face {X[~free] = 0, X 1 = 1}: per column
    n, X[I_n, n] = K_II^-1 (B[I_n, n] - lambda[I_n]); the row sums fix lambda
    through the Schur complement. Two steps of iterative refinement on the
    residuals of both equations recover the accuracy the explicit inverses of an
    ill-conditioned K lose."""
    inv, S = _face_factors(K, free)
    M = B.shape[0]
    X, lam = _face_solve(inv, S, B, np.ones(M))
    for _ in range(2):
        R = np.where(free, B - K @ X - lam[:, None], 0.0)
        dX, dlam = _face_solve(inv, S, R, 1.0 - X.sum(1))
        X += dX
        lam += dlam
    return X, lam


def exact(F, P, gamma: float, X0=None, maxit: int = 200):
    """Minimiser by a primal-dual active-set method with dense face solves (small
    problems): active where mu - d X > 0 (d = diag K). Returns (X, converged);
    converged means the active set repeated and X passes the KKT certificate."""
    K = hessian_block(F, gamma).toarray()
    B = 2.0 * (sps.csr_matrix(F).T @ np.asarray(P))
    M, N = B.shape
    X = np.full((M, N), 1.0 / N) if X0 is None else np.array(X0, dtype=np.float64)
    d = np.diag(K)[:, None]
    act, lam = None, None
    for _ in range(maxit):
        g = K @ X - B
        if lam is None:
            free_now = X > 0
            lam = np.array([-(g[i][free_now[i]].mean()) if free_now[i].any() else -g[i].min() for i in range(M)])
        mu = g + lam[:, None]
        new = (mu - d * X) > 0
        for i in np.nonzero(new.all(axis=1))[0]:
            new[i, np.argmax(X[i])] = False  # a row keeps at least one free entry
        if act is not None and np.array_equal(new, act):
            break
        act = new
        X, lam = _face_solution(K, B, ~act)
    X = np.where(X < 0.0, 0.0, X) if X.min() > -1e-13 else X
    return X, is_minimiser(F, P, gamma, X)


def brute_force(F, P, gamma: float):
    """Minimiser by enumerating every active set (tiny problems only, M*N <= 12):
    the face solution that is feasible and has the least objective."""
    K = hessian_block(F, gamma).toarray()
    B = 2.0 * (sps.csr_matrix(F).T @ np.asarray(P))
    M, N = B.shape
    best, bestf = None, np.inf
    for bits in itertools.product([False, True], repeat=M * N):
        free = np.array(bits).reshape(M, N)
        if not free.any(axis=1).all():
            continue
        try:
            X, _ = _face_solution(K, B, free)
        except np.linalg.LinAlgError:
            continue
        if X.min() < -1e-12:
            continue
        f = objective(F, P, gamma, X)
        if f < bestf:
            best, bestf = X, f
    return best


def face_polish(F, P, gamma: float, X0, tol: float = 1e-10, max_outer: int = 60, passes: int = 2,
                theta: float = 1e-13, globalise: bool = False, memory: int = 5, with_iterations: bool = False):
    """The direct polish's iteration (polish method 1 of the library), restated densely:
    primal-dual active-set steps in the full space with the row sums as constraints.
    The active set of a step is {G + lambda - d X > -theta d} (d = diag K), a row whose
    entries would all be active keeping its largest entry free; lambda starts at -G at
    each row's largest entry. Each face is solved by `passes` corrections from the
    current point with the column-block inverses and the Schur complement of the row
    sums, solving for the multiplier increment. Returns (X, status): 1 when the active
    set repeats after a last correction of at most 100 tol, 2 at the iteration cap
    (and the iteration count too with `with_iterations`).

    globalise: the face solution is taken as a direction instead of the next point, and
    the step is halved until the complementarity residual ||min(mu, d X)/d||^2 +
    ||1 - X 1||^2 falls below the largest of its last `memory` values (a nonmonotone
    Armijo test, sigma = 1e-4); certified when the active set repeats with that residual
    below 1e-24. Not in the library."""
    K = hessian_block(F, gamma).toarray()
    B = 2.0 * (sps.csr_matrix(F).T @ np.asarray(P))
    X = np.array(X0, dtype=np.float64)
    M, N = X.shape
    rows = np.arange(M)
    d = np.diag(K)[:, None]
    g = K @ X - B
    lam = -g[rows, X.argmax(1)]

    def resid(X, g, lam):
        r1 = np.minimum(g + lam[:, None], d * X) / d
        r2 = 1.0 - X.sum(1)
        return float((r1 ** 2).sum() + (r2 ** 2).sum())

    def out(X, status, it):
        return (X, status, it) if with_iterations else (X, status)

    act, step, hist = None, np.inf, []
    for it in range(max_outer + 1):
        new = (g + lam[:, None]) - d * X > -theta * d
        full_rows = new.all(1)
        new[rows[full_rows], X.argmax(1)[full_rows]] = False
        r0 = resid(X, g, lam) if globalise else 0.0
        done = step <= 100.0 * tol if not globalise else r0 < 1e-24
        if act is not None and np.array_equal(new, act) and done:
            return out(X, 1, it)
        if it == max_outer:
            return out(X, 2, it)
        act = new
        Xf = X.copy()
        Xf[act] = 0.0
        lf = lam.copy()
        inv, S = _face_factors(K, ~act)
        L = np.linalg.cholesky(S)
        for _ in range(passes):
            gf = K @ Xf - B
            Y = np.zeros((M, N))
            gl = gf + lf[:, None]
            for n, (I, Ki) in enumerate(inv):
                if I.size:
                    Y[I, n] = Ki @ gl[I, n]
            dl = np.linalg.solve(L.T, np.linalg.solve(L, -(1.0 - Xf.sum(1)) - Y.sum(1)))
            lf = lf + dl
            dX = np.zeros((M, N))
            for n, (I, Ki) in enumerate(inv):
                if I.size:
                    dX[I, n] = -(Y[I, n] + Ki @ dl[I])
            Xf += dX
            step = float(np.abs(dX).max())
        gf = K @ Xf - B
        if not globalise:
            X, g, lam = Xf, gf, lf
            continue
        hist.append(r0)
        ref = max(hist[-memory:])
        t = 1.0
        while True:
            Xt, gt, lt = X + t * (Xf - X), g + t * (gf - g), lam + t * (lf - lam)
            if resid(Xt, gt, lt) <= (1.0 - 2e-4 * t) * ref or t < 1e-3:
                break
            t *= 0.5
        X, g, lam = Xt, gt, lt
    return out(X, 2, max_outer)
