"""pqdts_amd.profsum on a synthetic kernel table (CPU only)."""
import sqlite3

from pqdts_amd import profsum


def test_lanes_busy_span_and_gaps(tmp_path, capsys):
    db = str(tmp_path / "k.db")
    c = sqlite3.connect(db)
    c.execute("create table kernels (name text, stream_id integer, start integer, duration integer)")
    ms = 1_000_000
    # stream 1: kernels at 0-2 ms and 5-6 ms (a 3 ms gap); stream 2: one kernel 1-9 ms
    c.executemany("insert into kernels values (?, ?, ?, ?)",
                  [("a", 1, 0, 2 * ms), ("b", 1, 5 * ms, 1 * ms), ("c", 2, 1 * ms, 8 * ms)])
    c.commit()
    profsum.lanes(db, 0.005)
    out = capsys.readouterr().out.splitlines()
    s1 = next(ln.split() for ln in out if ln.split()[:1] == ["1"])
    s2 = next(ln.split() for ln in out if ln.split()[:1] == ["2"])
    # stream, kernels, busy_ms, span_ms, busy%, gaps_ms, gaps>1ms, max_gap_ms
    assert s1[1:] == ["2", "3.00", "6.00", "50.0", "3.00", "1", "3.00"]
    assert s2[1:] == ["1", "8.00", "8.00", "100.0", "0.00", "0", "0.00"]
    # windows of 5 ms from t = 0: stream 1 busy 2 of 5 ms, then 1 of 5; stream 2 4 of 5, then 4 of 5
    rows = [ln.split() for ln in out if ln.split() and ln.split()[0] in ("0.000", "0.005")]
    assert rows == [["0.000", "0.40", "0.80"], ["0.005", "0.20", "0.80"]]
