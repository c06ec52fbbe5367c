"""GPU parity of the PIKG-shaped soft-force kernels against the oracle.

Bar: FP32 neighbour number/id_min/id_max bit-identical to the oracle's FP32 kernel (same
fixed operation order); FP32 acc/phi within the stated rms relative error of the FP64
oracle; the FP64 GPU twin within 1e-12 of the FP64 oracle."""
import numpy as np
import pytest

from gplum_amd import abi, gpu
from tests import helpers

pytestmark = pytest.mark.gpu

ACC_RMS_TOL = 1e-6     # FP32 kernel vs FP64, rms relative |acc| error
PHI_RMS_TOL = 1e-6


@pytest.fixture(scope="module")
def disk_tree(oracle_lib):
    p, pp = helpers.disk_with_radii(oracle_lib, 40000, seed=1)
    tree = oracle_lib.tree_build(p, pp)
    yield p, pp, tree
    tree.free()


@pytest.mark.parametrize("which", ["first", "middle", "last"])
def test_epep_epsp_group(oracle_lib, disk_tree, which):
    p, pp, tree = disk_tree
    g = {"first": 0, "middle": tree.n_group // 2, "last": tree.n_group - 1}[which]
    epi, epj, spq, spm = helpers.group_lists(oracle_lib, p, pp, tree, g)
    print("group", g, "ni", len(epi), "nj", len(epj), "ns", len(spq), flush=True)
    f_gpu = abi.cleared_force(len(epi))
    gpu.calc_force_epep(epi, epj, f_gpu)
    f_orc = abi.cleared_force(len(epi))
    oracle_lib.epep_f32(epi, epj, f_orc)
    assert np.array_equal(f_gpu["neighbor"], f_orc["neighbor"])
    gpu.calc_force_epsp_quad(epi, spq, f_gpu)
    f64 = abi.cleared_force(len(epi))
    a1, p1 = oracle_lib.epep_f64(epi, epj, force=f64)
    a2, p2 = oracle_lib.epsp_quad_f64(epi, spq)
    err_a = helpers.rel_err_rms(f_gpu["acc"], a1 + a2)
    err_p = helpers.rel_err_rms(f_gpu["phi"], p1 + p2)
    print("rms rel err acc %.3e phi %.3e" % (err_a, err_p))
    assert err_a < ACC_RMS_TOL and err_p < PHI_RMS_TOL
    # FP64 twins
    g64 = abi.cleared_force(len(epi))
    ga1, gp1 = gpu.calc_force_epep_f64(epi, epj, g64)
    ga2, gp2 = gpu.calc_force_epsp_quad_f64(epi, spq, g64)
    assert np.array_equal(g64["neighbor"], f64["neighbor"])
    assert helpers.rel_err_rms(ga1 + ga2, a1 + a2) < 1e-12
    assert helpers.rel_err_rms(gp1 + gp2, p1 + p2) < 1e-12
    # monopole
    fm = abi.cleared_force(len(epi))
    gpu.calc_force_epsp_mono(epi, spm, fm)
    am, pm = oracle_lib.epsp_mono_f64(epi, spm)
    assert helpers.rel_err_rms(fm["acc"], am) < 1e-5


def test_neighbour_predicate_at_threshold(oracle_lib):
    """j placed within +-1e-6 relative of the 1.0201 r_search^2 threshold: the GPU's FP32
    neighbour count equals the oracle's bit for bit (same operation order)."""
    rng = np.random.default_rng(3)
    ni, nj = 200, 3000
    epi = np.zeros(ni, dtype=abi.EPI)
    epi["id_local"] = np.arange(ni)
    epi["pos"] = 1.0 + rng.normal(0, 1e-4, (ni, 3))
    epi["r_out"] = 1e-4
    epi["r_search"] = 2e-4
    epj = np.zeros(nj, dtype=abi.EPJ)
    epj["id_local"] = ni + np.arange(nj)
    host = rng.integers(0, ni, nj)
    d = rng.normal(0, 1, (nj, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rad = 2e-4 * np.sqrt(1.0201) * (1 + rng.uniform(-1e-6, 1e-6, nj))
    epj["pos"] = epi["pos"][host] + d * rad[:, None]
    epj["r_out"] = 1e-4
    epj["r_search"] = 2e-4 * (1 + rng.uniform(-1e-7, 1e-7, nj))
    epj["mass"] = 1e-12
    f_gpu = abi.cleared_force(ni)
    gpu.calc_force_epep(epi, epj, f_gpu)
    f_orc = abi.cleared_force(ni)
    oracle_lib.epep_f32(epi, epj, f_orc)
    assert f_orc["neighbor"]["number"].sum() > 100
    assert np.array_equal(f_gpu["neighbor"], f_orc["neighbor"])


@pytest.mark.parametrize("ni,nj", [(1, 0), (1, 1), (63, 65), (257, 300), (1000, 17)])
def test_ragged_sizes(oracle_lib, ni, nj):
    rng = np.random.default_rng(ni * 7 + nj)
    pp = np.zeros(max(ni, nj), dtype=abi.PARTICLE)
    pp["pos"] = 1.0 + rng.normal(0, 1e-3, (len(pp), 3))
    pp["r_out"] = 2e-4
    pp["r_search"] = 3e-4
    pp["mass"] = 1e-10
    pp["id"] = np.arange(len(pp))
    epi = helpers.to_epi(pp, np.arange(ni))
    epj = helpers.to_epj(pp, np.arange(nj))
    f_gpu = abi.cleared_force(ni)
    gpu.calc_force_epep(epi, epj, f_gpu)
    f_orc = abi.cleared_force(ni)
    oracle_lib.epep_f32(epi, epj, f_orc)
    assert np.array_equal(f_gpu["neighbor"], f_orc["neighbor"])
    if nj:
        a, ph = oracle_lib.epep_f64(epi, epj)
        # a particle whose only j is itself feels no force: compare absolutely there
        na = np.linalg.norm(a, axis=1)
        err = np.linalg.norm(f_gpu["acc"] - a, axis=1)
        assert np.all(err[na == 0] == 0)
        rel = err[na > 0] / na[na > 0]
        if len(rel):
            print("max rel |acc| err %.3e" % rel.max())
            assert rel.max() < 1e-5
        assert np.abs(f_gpu["phi"] - ph).max() < 1e-5 * np.abs(ph).max()


def test_epsp_quadrupole_scaling_on_gpu(oracle_lib):
    """The GPU EP-SP kernels against the direct sum of a compact group's constituents
    (tests/test_oracle.py): the FP64 twin's quadrupole error falls as (s/d)^3, and the
    FP32 monopole/quadrupole kernels carry the same expansion error as the FP64 oracle."""
    from tests.test_oracle import epsp_errors
    dists = np.array([10.0, 20.0, 40.0, 80.0])

    def gpu_quad64(epi, sp):
        return gpu.calc_force_epsp_quad_f64(epi, sp, abi.cleared_force(len(epi)))

    def gpu32(fn):
        def call(epi, sp):
            f = abi.cleared_force(len(epi))
            fn(epi, sp, f)
            return f["acc"].astype(np.float64), f["phi"].astype(np.float64)
        return call

    e = epsp_errors(oracle_lib.epsp_mono_f64, gpu_quad64, np.random.default_rng(5), 1.0, dists)
    for col in range(2):
        slope_q = np.polyfit(np.log(dists), np.log(e[:, 1, col]), 1)[0]
        assert -3.3 < slope_q < -2.7, (col, slope_q)
    ref = epsp_errors(oracle_lib.epsp_mono_f64, oracle_lib.epsp_quad_f64, np.random.default_rng(5), 1.0, dists)
    np.testing.assert_allclose(e[:, 1, :], ref[:, 1, :], rtol=1e-6)
    e32 = epsp_errors(gpu32(gpu.calc_force_epsp_mono), gpu32(gpu.calc_force_epsp_quad),
                      np.random.default_rng(5), 1.0, dists[:3])
    np.testing.assert_allclose(e32, ref[:3], rtol=0.05)
