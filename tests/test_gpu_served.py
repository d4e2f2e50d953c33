"""The served form: the library's request server (pqdts_server_*).

A short arrival stream of requests from the uneven mix plus one large request
(M=8000, N=4000, D=16000, the served stream's large size, capped at 2 Newton
iterations per stage), in FP64 and mixed precision, through three size-class
lanes under a device-byte budget smaller than the stream's total footprint:
* every result is bit-identical to the same problem solved alone (FP64: the plain
  pqdts_solve; mixed: a batch of one), whatever it was batched with;
* the device bytes held by admitted problems never exceed the budget, requests
  waited in the queue for it, and every byte is returned at the end;
* pqdts_footprint's count is exactly what the context then holds.
"""
import time

import numpy as np
import pytest

import pqdts_amd
from pqdts_amd import workloads

pytestmark = pytest.mark.gpu

BOUNDS = (20000, 400000)


@pytest.fixture(scope="module")
def stream():
    mix = workloads.batch_mix(256, eps=1e-7)
    small = [mix[i] for i in (4, 13, 34, 35, 52, 94, 103, 206)]
    large = workloads.random_problem(8000, 4000, 16000, 0.8, 1e-3, 1e8, seed=900, eps=1e-7, name="large")
    return small, large


def _alone(p, maxiter, precision):
    c = pqdts_amd.Context.from_problem(p)
    c.pset(1.0 / p.N, "X")
    if precision == 0:
        conv, _ = c.solve(2, maxiter, p.eps, 1, 0)
    else:
        b = pqdts_amd.Batch(1)
        b.admit(c, 2, maxiter, p.eps, 1, 0, 1, record=False)
        conv = b.run()[0]["conv"]
        b.close()
    X = c.get("X")
    info = c.info()
    c.close()
    return X, conv, info


def test_footprint_is_what_the_context_holds(stream):
    small, large = stream
    for p in (small[0], small[3], large):
        for prec in (0, 1):
            pers, trans = pqdts_amd.footprint(p, prec)
            c = pqdts_amd.Context.from_problem(p)
            if prec:
                b = pqdts_amd.Batch(1)  # admission in mixed precision adds the FP32 copies
                b.admit(c, 2, 1, p.eps, 1, 0, 1, record=False)
                held = c.info()["device_bytes"]
                b.close()
            else:
                held = c.info()["device_bytes"]
            assert held == pers, (p.name, prec, held, pers)
            # the Gram band is formed in place: only a temporary descriptor on top
            if c.info()["gram"]:
                assert 0 < trans < 4096
            else:
                assert trans == 0
            c.close()


def test_served_stream_matches_alone_and_respects_budget(stream):
    small, large = stream
    reqs = [(p, 300, 0) for p in small[:5]] + [(large, 2, 0)] + [(p, 300, 1) for p in small[5:]] + \
           [(small[0], 300, 1)]
    fps = [sum(pqdts_amd.footprint(p, prec)) for p, _, prec in reqs]
    # room for the large request and a few small ones at a time, not for everything
    budget = max(fps) + 3 * sorted(fps)[-2]
    assert budget < sum(fps)
    srv = pqdts_amd.Server(BOUNDS, capacity=8, budget_bytes=budget)
    ids = {}
    for k, (p, maxiter, prec) in enumerate(reqs):
        ids[srv.submit(p, maxiter=maxiter, precision=prec)] = k
        time.sleep(0.02)
    out = {}
    t0 = time.time()
    while len(out) < len(reqs):
        for rid in srv.wait(timeout_ms=1000):
            out[ids[rid]] = srv.result(rid)
        assert srv.info()["peak_bytes"] <= budget
        assert time.time() - t0 < 300
    info = srv.info()
    srv.close()
    assert info["peak_bytes"] <= budget and info["bytes_in_use"] == 0
    assert info["completed"] == len(reqs) == info["submitted"]
    assert max(r["queued_s"] for r in out.values()) > 0.0
    for k, (p, maxiter, prec) in enumerate(reqs):
        Xa, conva, _ = _alone(p, maxiter, prec)
        r = out[k]
        assert r["finished"]
        assert np.array_equal(r["X"], Xa), f"request {k} ({p.name}, precision {prec})"
        assert r["conv"] == conva


def test_served_polish_matches_alone(stream):
    """Polished requests (server-wide polish settings, FP64 and mixed with FP32
    polish products) return the bits of the same polished solve alone, and the
    polish outcome with them."""
    small, _ = stream
    reqs = [(small[k], prec) for k in (1, 3, 6) for prec in (0, 1)]
    srv = pqdts_amd.Server(BOUNDS, capacity=8)
    srv.set_polish(40, fp32=True)
    ids = {srv.submit(p, maxiter=300, precision=prec): k for k, (p, prec) in enumerate(reqs)}
    out = {}
    t0 = time.time()
    while len(out) < len(reqs):
        for rid in srv.wait(timeout_ms=1000):
            out[ids[rid]] = srv.result(rid)
        assert time.time() - t0 < 300
    srv.close()
    for k, (p, prec) in enumerate(reqs):
        c = pqdts_amd.Context.from_problem(p)
        c.set_polish(40, fp32=True)
        c.pset(1.0 / p.N, "X")
        if prec == 0:
            c.solve(2, 300, p.eps, 1, 0)
        else:
            b = pqdts_amd.Batch(1)
            b.admit(c, 2, 300, p.eps, 1, 0, 1, record=False)
            b.run()
            b.release(0)
            b.close()
        pol = c.polish_result()
        assert np.array_equal(out[k]["X"], c.get("X")), (p.name, prec)
        assert out[k]["polish"] == {"status": pol["status"], "iterations": pol["iterations"],
                                    "cg_rounds": pol["cg_rounds"]}, (p.name, prec)
        c.close()
