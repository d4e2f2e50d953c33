"""CPU checks of the native boundary: the C-ABI library loads and exports every
symbol include/pqdts_amd.h declares; host-side formats round-trip. No GPU calls."""
import ctypes
import os
import re

import numpy as np

import pqdts_amd
from pqdts_amd import _lib, fortio, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "pqdts_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:pqdts|simplexproj)_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/pqdts_amd.h but not exported"
    assert sorted(_lib.EXPORTS) == syms


def test_python_binding_loads():
    L = pqdts_amd.load()
    assert L is not None


def test_state_file_roundtrip(tmp_path):
    X = np.random.default_rng(0).random((7, 3))
    p = tmp_path / "s.dat"
    fortio.write_state(str(p), 7, 3, 11, X)
    M, N, D, lr, rank, Y = fortio.read_state(str(p))
    assert (M, N, D, lr, rank) == (7, 3, 11, 7, 0)
    assert np.array_equal(X, Y)
    # record layout of the reference: 4 x (4+8+4) + (4+4+4) + (4 + 8*M*N + 4)
    assert p.stat().st_size == 4 * 16 + 12 + 8 + 8 * 21


def test_coo_files_roundtrip(tmp_path):
    prob = workloads.config1()
    NF, NP = workloads.write_inputs(str(tmp_path), prob)
    assert NF == prob.F.nnz
    r, c, v = workloads.read_coo(str(tmp_path / "data"), "F", NF)
    assert r.dtype == np.int32 and c.dtype == np.int32 and v.dtype == np.float64
    F = np.zeros((prob.D, prob.M))
    F[r, c] = v
    assert np.array_equal(F, prob.F.toarray())


def test_cu_mask_must_be_one_consecutive_range():
    """pqdts_batch_set_cu_mask checks the mask before anything else, with no GPU
    call: masks whose CUs do not form one consecutive range are refused (a host
    crash was once observed while such a stream was being set up)."""
    L = pqdts_amd.load()
    words, n = _lib.cu_mask_words([0, 1, 2, 8, 9], 256)
    assert n == 9
    assert L.pqdts_batch_set_cu_mask(None, words, n) != 0
    assert b"consecutive" in L.pqdts_last_error()
    words, n = _lib.cu_mask_words(range(0, 256, 8), 256)
    assert L.pqdts_batch_set_cu_mask(None, words, n) != 0
    assert b"consecutive" in L.pqdts_last_error()
    words, n = _lib.cu_mask_words([], 256)
    assert L.pqdts_batch_set_cu_mask(None, words, n) != 0
    assert b"no CU" in L.pqdts_last_error()
    # a consecutive range passes the mask check and only then meets the null batch
    words, n = _lib.cu_mask_words(range(32, 64), 256)
    assert L.pqdts_batch_set_cu_mask(None, words, n) != 0
    assert b"null batch" in L.pqdts_last_error()


def test_tuned_split_table_is_well_formed():
    """The autotuner's committed table (read by the library's split rule): 11 x 10
    buckets, each 0 (no tuned shape: the rule) or a split of 1, 2 or 4."""
    txt = open(os.path.join(ROOT, "pqdts_amd", "csrc", "tuned_split.inc")).read()
    body = txt[txt.index("= {") + 3: txt.rindex("};")]
    rows = re.findall(r"\{([^}]*)\}", body)
    assert len(rows) == 11
    for r in rows:
        vals = [int(v) for v in r.split(",")]
        assert len(vals) == 10 and set(vals) <= {0, 1, 2, 4}


def test_footprint_runs_on_the_host():
    """pqdts_footprint needs no GPU: it plans the context from F's pattern. Mixed
    precision adds the FP32 band copies and the FP32 direction copy (more than 4 M N
    bytes); a problem on the Gram path needs only a descriptor on top while its band
    is formed; a wider Pi costs at least its six work arrays per extra column."""
    prob = workloads.config1()
    p64, t64 = pqdts_amd.footprint(prob, 0)
    p32, t32 = pqdts_amd.footprint(prob, 1)
    assert p64 > 8 * 6 * prob.M * prob.N and p32 > p64 + 4 * prob.M * prob.N
    assert 0 <= t64 < 4096 and t64 == t32
    import dataclasses
    wide = dataclasses.replace(prob, N=2 * prob.N, P=np.zeros((prob.D, 2 * prob.N)))
    assert pqdts_amd.footprint(wide, 0)[0] >= p64 + 48 * prob.M * prob.N
