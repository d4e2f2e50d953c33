"""Inputs of the drop-in binary cases that the reference build pins.

Each case writes its data files into a work directory and returns the 14
command-line arguments (pqdts.f90:1243-1270). make_golden.py runs the reference
binary on them and records what it printed and wrote; tests/test_gpu_driver.py
runs the new binary on the same inputs and compares. Every case stops inside the
part of the trajectory that the reference reproduces across thread counts (the
first Newton iteration of each stage; three projected-gradient iterations for
mode 1).

  intF        F generated inside the program (argument 4 = -1, M = (D-1)^2,
              pqdts.f90:1366-1466), P from a file
  smooth      restart from data/initial_     0.dat with smoothing (smo_fact = 20,
              pqdts.f90:1877-1898; only rows j > 100 are smoothed, hence M = 200)
  mode1       projected gradient (mode 1, pqdts.f90:1934-1952)
  randomP     P drawn by the program (argument 5 = -1, pqdts.f90:1537)
  stage2      start in stage 2 (pqdts.f90:1965-1983) from a restart state. From the
              default start Pi = 1/N the stage-2 transform's maxloc is decided by
              the last bit of a row sum of N copies of 1/N: which entry ends up
              largest depends on the order the compiler sums the row in (the random
              maxr of pqdts.f90:1565-1575 only enters through that rounding), so no
              other summation order reproduces it. From a row-stochastic state
              with distinct entries the argmax is determined.
"""
import os

import numpy as np
import scipy.sparse as sps

from pqdts_amd import fortio, workloads

CASES = ("intF", "smooth", "mode1", "randomP", "stage2")


def _args(M, N, D, NF, NP, mode=2, maxiter=1, output=1, gamma=1e-4, eps=1e-6, start_stage=1, read_input=0,
          smo=0.0, bench=0):
    return [M, N, D, NF, NP, mode, maxiter, output, gamma, eps, start_stage, read_input, smo, bench]


def build(name, workdir):
    """Writes the inputs of case `name` under workdir; returns the 14 arguments."""
    data = os.path.join(workdir, "data")
    os.makedirs(data, exist_ok=True)
    if name == "intF":
        D, N = 12, 6
        M = (D - 1) ** 2
        P = np.random.default_rng(0).random((D, N))
        P /= P.sum(1, keepdims=True)
        NP = workloads.write_coo(data, "P", sps.csr_matrix(P))
        return _args(M, N, D, -1, NP, gamma=1e-3)
    if name == "smooth":
        prob = workloads.random_problem(200, 60, 400, 0.8, 1e-3, 1e6, seed=3, eps=1e-6, name="smooth")
        NF, NP = workloads.write_inputs(workdir, prob)
        X0 = 0.5 * prob.pi_true + 0.5 / prob.N
        fortio.write_state(os.path.join(data, "initial_     0.dat"), prob.M, prob.N, prob.D, X0, rank_bytes=8)
        return _args(prob.M, prob.N, prob.D, NF, NP, gamma=prob.gamma, read_input=1, smo=20.0)
    prob = workloads.config1()
    if name == "mode1":
        NF, NP = workloads.write_inputs(workdir, prob)
        return _args(prob.M, prob.N, prob.D, NF, NP, mode=1, maxiter=3, gamma=prob.gamma)
    if name == "randomP":
        NF = workloads.write_coo(data, "F", prob.F)
        return _args(prob.M, prob.N, prob.D, NF, -1, gamma=prob.gamma)
    if name == "stage2":
        NF, NP = workloads.write_inputs(workdir, prob)
        X0 = 0.5 * prob.pi_true + 0.5 / prob.N
        fortio.write_state(os.path.join(data, "initial_     0.dat"), prob.M, prob.N, prob.D, X0, rank_bytes=8)
        return _args(prob.M, prob.N, prob.D, NF, NP, gamma=prob.gamma, start_stage=2, read_input=1)
    raise KeyError(name)
