"""Snapshot I/O in the reference's ASCII format (host code of the library, no GPU):
the reference's own sample snapshot reads back and writes out byte for byte."""
import gzip
import os

import numpy as np

from gplum_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_init3000_round_trip(tmp_path, init3000):
    src = tmp_path / "INIT3000.dat"
    with gzip.open(os.path.join(ROOT, "tests", "golden", "INIT3000.dat.gz"), "rb") as fh:
        raw = fh.read()
    src.write_bytes(raw)
    h, pp = gpu.snapshot_read(str(src))
    hdr, ref = init3000
    assert h.n_body == len(pp) == 3000 and h.id_next == 3000
    assert h.time == 0.0
    assert list(h.e_init) == list(hdr[3:8])
    assert np.array_equal(pp["id"], np.arange(3000))
    assert np.array_equal(pp["pos"], ref["pos"]) and np.array_equal(pp["vel"], ref["vel"])
    assert np.array_equal(pp["mass"], ref["mass"]) and np.array_equal(pp["r_planet"], ref["r_planet"])
    out = tmp_path / "out.dat"
    gpu.snapshot_write(str(out), h, pp)
    assert out.read_bytes() == raw


def test_missing_file_is_an_error(tmp_path):
    try:
        gpu.snapshot_read(str(tmp_path / "nope.dat"))
    except gpu.GplumError as e:
        assert "cannot open" in str(e)
    else:
        raise AssertionError("no error for a missing file")
