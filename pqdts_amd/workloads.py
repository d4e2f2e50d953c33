"""Synthetic detector-tomography workloads.

Every problem is  min_Pi ||P - F Pi||^2 + gamma * sum_n sum_i (Pi[i,n] - Pi[i+1,n])^2
with rows of Pi on the probability simplex (pqdts README.md:4-16).

* F (D x M): F[d, i] = Poisson(i; alpha_d^2), the photon-number statistics of a
  coherent probe, entries below 1e-14 dropped, stored as CSR.
* Pi_true (M x N): the detector's true POVM, built from a loss channel
  (binomial thinning with efficiency eta) followed by a read-out map
  (saturation at N-1, or a multiplexed-pixel occupancy read-out).
* P (D x N): each row of F Pi_true normalised and resampled multinomially with
  the stated number of shots, one draw per row in row order from
  numpy.random.default_rng(seed).

The solver never sees Pi_true; tests use it for sanity bounds only.
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np
import scipy.sparse as sps
from scipy.special import gammaln

POISSON_CUT = 1e-14


@dataclasses.dataclass
class Problem:
    name: str
    M: int
    N: int
    D: int
    F: sps.csr_matrix      # D x M, float64
    P: np.ndarray          # D x N, float64, C order
    gamma: float
    eps: float
    pi_true: np.ndarray | None = None
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.F.nnz)


def poisson_rows(alpha2: np.ndarray, M: int, cut: float = POISSON_CUT) -> sps.csr_matrix:
    """F[d, i] = exp(i log a - a - log i!) for i = 0..M-1, dropping values < cut."""
    i = np.arange(M, dtype=np.float64)
    lgf = gammaln(i + 1.0)
    rows, cols, vals = [], [], []
    for d, a in enumerate(alpha2):
        if a == 0.0:
            pmf = np.zeros(M)
            pmf[0] = 1.0
        else:
            # only evaluate a window around the mean; outside it the pmf is < cut
            half = int(12.0 * math.sqrt(a) + 40)
            lo = max(0, int(a) - half)
            hi = min(M, int(a) + half + 1)
            pmf = np.zeros(M)
            ii = i[lo:hi]
            pmf[lo:hi] = np.exp(ii * math.log(a) - a - lgf[lo:hi])
        keep = np.nonzero(pmf >= cut)[0]
        rows.append(np.full(keep.size, d, dtype=np.int64))
        cols.append(keep)
        vals.append(pmf[keep])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    return sps.csr_matrix((v, (r, c)), shape=(len(alpha2), M))


def loss_matrix(M: int, eta: float, cut: float = 0.0) -> sps.csr_matrix:
    """B[i, k] = Binom(k; i, eta): i photons, each transmitted with probability eta."""
    if eta >= 1.0:
        return sps.identity(M, format="csr")
    i = np.arange(M, dtype=np.float64)[:, None]
    k = np.arange(M, dtype=np.float64)[None, :]
    with np.errstate(invalid="ignore", divide="ignore"):
        logp = (gammaln(i + 1) - gammaln(k + 1) - gammaln(np.maximum(i - k, 0) + 1)
                + k * math.log(eta) + (i - k) * math.log1p(-eta))
    B = np.where(k <= i, np.exp(logp), 0.0)
    B[B <= cut] = 0.0
    return sps.csr_matrix(B)


def saturate(K: int, N: int) -> sps.csr_matrix:
    """S[k, n] = [min(k, N-1) == n]."""
    k = np.arange(K)
    return sps.csr_matrix((np.ones(K), (k, np.minimum(k, N - 1))), shape=(K, N))


def occupancy(K: int, npix: int) -> np.ndarray:
    """Occ[k, j] = P(j of npix pixels lit | k photons spread uniformly).

    p_k(j) = p_{k-1}(j) j/npix + p_{k-1}(j-1) (npix-j+1)/npix, p_0(0) = 1.
    """
    occ = np.zeros((K, npix + 1))
    cur = np.zeros(npix + 1)
    cur[0] = 1.0
    j = np.arange(npix + 1, dtype=np.float64)
    for k in range(K):
        occ[k] = cur
        nxt = cur * j / npix
        nxt[1:] += cur[:-1] * (npix - j[1:] + 1.0) / npix
        cur = nxt
    return occ


def sample_P(FPi: np.ndarray, shots: float | np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    D = FPi.shape[0]
    shots = np.broadcast_to(np.asarray(shots, dtype=np.float64), (D,))
    P = np.empty_like(FPi)
    for d in range(D):
        q = np.clip(FPi[d], 0.0, None)
        q = q / q.sum()
        s = int(shots[d])
        P[d] = rng.multinomial(s, q) / s
    return np.ascontiguousarray(P)


def config1(seed: int = 0) -> Problem:
    """M=100, N=50, D=200, alpha^2_d = 80 d/(D-1), eta=0.8, saturation at N-1, 1e6 shots,
    gamma=1e-4, eps=1e-6 (the reference's CPU-sized case)."""
    M, N, D = 100, 50, 200
    alpha2 = 80.0 * np.arange(D) / (D - 1)
    F = poisson_rows(alpha2, M)
    pi = (loss_matrix(M, 0.8) @ saturate(M, N)).toarray()
    P = sample_P(F @ pi, 1e6, seed)
    return Problem("cfg1", M, N, D, F, P, 1e-4, 1e-6, pi, {"eta": 0.8, "shots": 1e6})


def config2(seed: int = 0) -> Problem:
    """M=2000, N=1000, D=4000, alpha^2 = linspace(0, 1800, D), eta=0.75, 1000-pixel
    multiplexed read-out saturating at N-1, 1e7 shots, gamma=1e-3, eps=1e-7."""
    M, N, D, npix = 2000, 1000, 4000, 1000
    alpha2 = np.linspace(0.0, 1800.0, D)
    F = poisson_rows(alpha2, M)
    occ = occupancy(M, npix)                          # M x (npix+1)
    pi = loss_matrix(M, 0.75, 1e-300) @ (sps.csr_matrix(occ) @ saturate(npix + 1, N))
    pi = np.asarray(pi.toarray() if sps.issparse(pi) else pi)
    P = sample_P(F @ pi, 1e7, seed)
    return Problem("cfg2", M, N, D, F, P, 1e-3, 1e-7, pi, {"eta": 0.75, "shots": 1e7, "npix": npix})


def config3(seed: int = 0) -> Problem:
    """M=20000, N=10000, D=40000, alpha^2 = linspace(0, 18000, D), eta=0.9, saturation at
    N-1, 1e8 shots, gamma=1e-2, eps=1e-8 (the single large instance)."""
    return random_problem(20000, 10000, 40000, 0.9, 1e-2, 1e8, seed, eps=1e-8, name="cfg3")


def thinned_outcomes(alpha2: np.ndarray, eta: float, N: int) -> np.ndarray:
    """Outcome distribution of a loss (efficiency eta) + saturation (at N-1) detector
    probed by coherent states: binomial thinning of Poisson(a) photons is
    Poisson(eta a), so q[d, n] = Poisson(n; eta a_d) for n < N-1 and the tail at N-1."""
    lam = eta * np.asarray(alpha2, dtype=np.float64)[:, None]
    n = np.arange(N - 1, dtype=np.float64)[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        logp = n * np.log(np.where(lam > 0, lam, 1.0)) - lam - gammaln(n + 1)
    q = np.exp(logp)
    q[:, :][lam[:, 0] == 0] = 0.0
    q[lam[:, 0] == 0, 0] = 1.0
    tail = np.clip(1.0 - q.sum(1), 0.0, None)
    return np.concatenate([q, tail[:, None]], axis=1)


def random_problem(M: int, N: int, D: int, eta: float, gamma: float, shots: float,
                   seed: int, eps: float = 1e-6, name: str = "rand") -> Problem:
    """Loss + saturation detector with alpha^2 = linspace(0, 0.9 M, D). P is drawn
    from the untruncated outcome distribution (thinned_outcomes), as a detector
    would record it; F models photon numbers below M only, so for the brightest
    probes F Pi_true misses the Poisson tail above M-1 that P contains."""
    alpha2 = np.linspace(0.0, 0.9 * M, D)
    F = poisson_rows(alpha2, M)
    pi = (loss_matrix(M, eta, 1e-300) @ saturate(M, N)).toarray() if M <= 2500 else None
    P = sample_P(thinned_outcomes(alpha2, eta, N), shots, seed)
    return Problem(name, M, N, D, F, P, gamma, eps, pi,
                   {"eta": eta, "shots": shots, "seed": seed})


def mix_specs(count: int = 256, seed: int = 4, m_lo: int = 50, m_hi: int = 2000,
              eps: float = 1e-6) -> list[tuple]:
    """The parameters of batch_mix's problems, in order (cheap: nothing is built)."""
    rng = np.random.default_rng(seed)
    specs = []
    for b in range(count):
        M = int(round(math.exp(rng.uniform(math.log(m_lo), math.log(m_hi)))))
        N = int(rng.integers(max(2, M // 4), M + 1))
        eta = float(rng.uniform(0.5, 0.95))
        gamma = float(math.exp(rng.uniform(math.log(1e-5), math.log(1e-2))))
        shots = float(round(math.exp(rng.uniform(math.log(1e6), math.log(1e8)))))
        specs.append((M, N, 2 * M, eta, gamma, shots, seed * 100003 + b, eps, f"mix{b}"))
    return specs


def batch_mix(count: int = 256, seed: int = 4, m_lo: int = 50, m_hi: int = 2000,
              eps: float = 1e-6, workers: int = 0, only=None) -> list[Problem]:
    """Uneven instances: M log-uniform in [m_lo, m_hi], N uniform in [M/4, M], D = 2M,
    eta uniform in [0.5, 0.95], gamma log-uniform in [1e-5, 1e-2], shots log-uniform
    in [1e6, 1e8]. only: build just these indices of the mix (same problems)."""
    specs = mix_specs(count, seed, m_lo, m_hi, eps)
    if only is not None:
        specs = [specs[i] for i in only]
    return build_many(specs, workers)


def _build(spec):
    M, N, D, eta, gamma, shots, seed, eps, name = spec
    return random_problem(M, N, D, eta, gamma, shots, seed, eps, name=name)


def build_many(specs, workers: int = 0) -> list[Problem]:
    """random_problem for every spec, in worker processes (each spec is seeded on its own).
    Workers are spawned (not forked), so this is safe after the GPU runtime started."""
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(specs) < 4:
        return [_build(s) for s in specs]
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_build, specs, chunksize=2))


# ---------------------------------------------------------------------------
# data/*.bin files as the reference reads them (pqdts.f90:1080-1107; the writer
# is pqdts.py:15-28): raw int32 row, int32 col (0-based), float64 values.
# ---------------------------------------------------------------------------

def write_coo(dirname: str, name: str, A) -> int:
    coo = sps.coo_matrix(A)
    os.makedirs(dirname, exist_ok=True)
    coo.row.astype(np.int32).tofile(os.path.join(dirname, f"{name}_row.bin"))
    coo.col.astype(np.int32).tofile(os.path.join(dirname, f"{name}_col.bin"))
    coo.data.astype(np.float64).tofile(os.path.join(dirname, f"{name}_data.bin"))
    return int(coo.nnz)


def write_inputs(workdir: str, prob: Problem, with_F: bool = True) -> tuple[int, int]:
    """Writes data/F_* and data/P_* under workdir; returns (NF, NP)."""
    d = os.path.join(workdir, "data")
    NF = write_coo(d, "F", prob.F) if with_F else -1
    NP = write_coo(d, "P", sps.csr_matrix(prob.P))
    return NF, NP


def read_coo(dirname: str, name: str, nnz: int):
    r = np.fromfile(os.path.join(dirname, f"{name}_row.bin"), dtype=np.int32, count=nnz)
    c = np.fromfile(os.path.join(dirname, f"{name}_col.bin"), dtype=np.int32, count=nnz)
    v = np.fromfile(os.path.join(dirname, f"{name}_data.bin"), dtype=np.float64, count=nnz)
    return r, c, v


CONFIGS = {"cfg1": config1, "cfg2": config2, "cfg3": config3}
