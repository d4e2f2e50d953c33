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
