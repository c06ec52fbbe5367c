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


def _random_state(init3000, seed=2):
    """INIT3000 particles with every field the restart file carries filled in."""
    from gplum_amd import abi
    rng = np.random.default_rng(seed)
    _, ref = init3000
    pp = np.zeros(len(ref), dtype=abi.PARTICLE)
    for f in ("id", "mass", "r_planet", "f", "pos", "vel"):
        pp[f] = ref[f]
    for f in ("acc", "acc_d", "jerk_d", "acc_s", "jerk_s"):
        pp[f] = rng.normal(size=(len(pp), 3))
    for f in ("phi", "phi_d", "phi_s", "r_out", "r_out_inv", "r_search", "v_disp", "time", "dt", "acc0"):
        pp[f] = rng.random(len(pp))
    pp["id_cluster"] = rng.integers(0, 3000, len(pp))
    pp["n_cluster"] = rng.integers(0, 50, len(pp))
    for f in ("number", "rank", "id_max", "id_min"):
        pp["neighbor"][f] = rng.integers(-5, 100, len(pp))
    pp["flags"] = rng.integers(0, 4, len(pp))
    return pp


def test_binary_restart_matches_oracle_layout(tmp_path, init3000, oracle_lib):
    """snap_tmp.dat: the library's flat records equal, byte for byte, the objects the
    oracle declares the way particle.h declares EPIGrav <- EPJGrav <- FPGrav (laid out by
    the C compiler), and both read back to the same particles."""
    import ctypes
    pp = _random_state(init3000)
    h = gpu.FileHeader()
    h.id_next, h.time = 3001, 0.75
    e_i = [-1.0, 0.5, -1.2, -0.3, -0.25, -0.05, 0.0]
    e_n = [-1.1, 0.6, -1.3, -0.4, -0.3, -0.1, 2e-9]
    for k, name in enumerate(gpu.ENERGY_FIELDS):
        setattr(h.e_init, name, e_i[k])
        setattr(h.e_now, name, e_n[k])
    lib_path = str(tmp_path / "snap_tmp_lib.dat")
    orc_path = str(tmp_path / "snap_tmp_orc.dat")
    gpu.restart_write(lib_path, h, pp)
    L = oracle_lib.lib()
    ei = np.array(e_i)
    en = np.array(e_n)
    assert L.orc_restart_write(orc_path.encode(), 3001, 0.75, ei.ctypes.data, en.ctypes.data,
                               pp.ctypes.data, len(pp)) == 0
    raw = open(lib_path, "rb").read()
    assert raw == open(orc_path, "rb").read()
    assert len(raw) == L.orc_file_header_size() + len(pp) * L.orc_fpgrav_size() == 128 + 344 * 3000
    # fields at their byte offsets (FPGrav: id at 48, mass at 56, id_cluster at 328)
    rec = np.frombuffer(raw[128:], dtype=np.uint8).reshape(len(pp), 344)
    assert np.array_equal(rec[:, 48:56].copy().view("<i8")[:, 0], pp["id"])
    assert np.array_equal(rec[:, 56:64].copy().view("<f8")[:, 0], pp["mass"])
    assert np.array_equal(rec[:, 328:336].copy().view("<i8")[:, 0], pp["id_cluster"])
    h2, back = gpu.restart_read(lib_path)
    assert (h2.n_body, h2.id_next, h2.time) == (3000, 3001, 0.75)
    assert [getattr(h2.e_now, n) for n in gpu.ENERGY_FIELDS] == e_n
    assert back.tobytes() == pp.tobytes()
    from gplum_amd import abi
    ob = np.zeros(len(pp), dtype=abi.PARTICLE)
    nb, idn, t = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_double(0)
    e1, e2 = np.zeros(7), np.zeros(7)
    assert L.orc_restart_read(lib_path.encode(), ctypes.byref(nb), ctypes.byref(idn), ctypes.byref(t),
                              e1.ctypes.data, e2.ctypes.data, ob.ctypes.data, len(ob)) == 0
    assert ob.tobytes() == pp.tobytes() and nb.value == 3000 and list(e2) == e_n
