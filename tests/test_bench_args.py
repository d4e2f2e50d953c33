"""Host logic of bench.py: the per-lane compute-unit layouts of --lane-cus."""
import os
import sys
import types

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

DEV = types.SimpleNamespace(device_cu_count=lambda: 256)


def parse(spec, lanes=3):
    return bench.parse_lane_cus(types.SimpleNamespace(lane_cus=spec, lanes=lanes), DEV)


def test_lane_cus_forms():
    assert parse("none") is None and parse("") is None
    assert bench.lane_cus_str(parse("auto")) == ["0-31", "all", "64-255"]
    assert parse("auto", lanes=2) is None
    assert bench.lane_cus_str(parse("32,64,160")) == ["0-31", "32-95", "96-255"]
    assert bench.lane_cus_str(parse("0-31,all,64-255")) == ["0-31", "all", "64-255"]
    assert list(parse("0-3,all,4-7")[0]) == [0, 1, 2, 3]


@pytest.mark.parametrize("spec", ["32,64", "32,64,200", "0-31,all,64-256", "0,all,1-2"])
def test_lane_cus_rejects(spec):
    with pytest.raises(SystemExit):
        parse(spec)


def test_polish_set_takes_the_smaller_size_classes():
    """--polish-classes k polishes the problems of the k smallest of the three size
    classes (thirds by M x N), the lanes' own split."""
    probs = [types.SimpleNamespace(M=m, N=n) for m, n in [(50, 20), (2000, 1000), (100, 90), (700, 300),
                                                          (60, 60), (1500, 400)]]
    assert bench.polish_set(probs, 0) == set()
    assert bench.polish_set(probs, 1) == {0, 4}
    assert bench.polish_set(probs, 2) == {0, 4, 2, 3}
    assert bench.polish_set(probs, 3) == set(range(6))


def test_certified_reads_the_polish_status():
    bench.POLISH = 60
    try:
        assert bench.certified({"polish": {"status": 1}})
        assert not bench.certified({"polish": {"status": 2}})
        assert not bench.certified({"polish": {"status": 0}})
        assert not bench.certified({})
    finally:
        bench.POLISH = 0
    assert bench.certified({})


def test_make_ctx_passes_the_polish_method():
    """The bench's contexts get the polish method of --polish-method (the direct face
    solves by default) and only where the polish applies."""
    calls = []

    class Ctx:
        @classmethod
        def from_problem(cls, prob):
            return cls()

        def set_polish(self, *a, **k):
            calls.append((a, k))

    fake = types.SimpleNamespace(Context=Ctx)
    saved = bench.POLISH, bench.POLISH_METHOD
    try:
        bench.POLISH, bench.POLISH_METHOD = 60, "direct"
        bench.make_ctx(fake, None, polish=True)
        bench.make_ctx(fake, None, polish=False)
        bench.POLISH_METHOD = "cg"
        bench.make_ctx(fake, None, polish=True)
    finally:
        bench.POLISH, bench.POLISH_METHOD = saved
    assert [k["method"] for _, k in calls] == ["direct", "cg"]
    assert all(a == (60,) for a, _ in calls)
    from pqdts_amd import _lib
    assert _lib.POLISH_METHODS == {"cg": 0, "direct": 1}
