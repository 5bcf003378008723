"""Solution files in the reference's text format (BVPDAEReadWriteData.py:14-96)."""
import numpy as np

from colloc_amd import bvpdae_io


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(0)
    N, ny, nz, np_ = 7, 4, 2, 3
    t = np.sort(rng.uniform(0, 3, N))
    y, z, p = rng.normal(size=(N, ny)), rng.normal(size=(N, nz)), rng.normal(size=np_) * 1e-7
    f = str(tmp_path / "sol.data")
    assert bvpdae_io.bvpdae_write_data(f, N, ny, nz, np_, t, y, z, p) == 0
    err, t2, y2, z2, p2 = bvpdae_io.bvpdae_read_data(f)
    assert err == 0
    for a, b in ((t, t2), (y, y2), (z, z2), (p, p2)):
        assert np.array_equal(a, b)


def test_layout(tmp_path):
    f = str(tmp_path / "small.data")
    bvpdae_io.bvpdae_write_data(f, 2, 1, 1, 1, np.array([0.0, 1.0]), np.array([[1.5], [2.5]]),
                                np.array([[3.0], [4.0]]), np.array([5.0]))
    lines = open(f).read().split("\n")
    assert lines[:4] == ["nt:", "2", "t:", "0.0 1.0 "]
    assert lines[4:9] == ["ny:", "1", "y:", "1.5 ", "2.5 "]
    assert lines[9:14] == ["nz:", "1", "z:", "3.0 ", "4.0 "]
    assert lines[14:17] == ["np:", "1", "p:"]


def test_unreadable_file_reports_error(tmp_path):
    assert bvpdae_io.bvpdae_read_data(str(tmp_path / "missing.data"))[0] == 1
    bad = tmp_path / "short.data"
    bad.write_text("nt:\n3\nt:\n0 1\n")
    assert bvpdae_io.bvpdae_read_data(str(bad))[0] == 1
