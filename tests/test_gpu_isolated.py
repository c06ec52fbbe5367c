"""GPU parity of the isolated-particle branch of HardSystem::timeIntegrate (hard.h:793-818),
the kernel k_isolated, against the oracle (orc_isolated), both branches:

* Sun-only Hermite (timeIntegrate_isolated, hermite.h:734-785), taken for e >= 0.8 or a
  softened Sun: only +, -, *, / and sqrt, so bit-identical to the oracle (positions,
  velocities, Sun and hard terms, time and block step);
* Kepler drift (timeIntegrateKepler_isolated, hermite.h:787-815) for e < 0.8: sin, cos
  and atan2 come from the device's and the host's libm, so within 1e-13 relative; its
  solveKeplerEq takes the fabs outcome of kepler.h:23 (the int abs(int) outcome is off by
  far more at e = 0.79).

The particles sit on orbits of e from 0.05 to 0.97 spread over 0.6-1.8 AU with Hill-sized
radii (R_cut1 = R_search1 = 0), so none has a neighbour and every one takes this branch."""
import numpy as np
import pytest

from gplum_amd import abi, gpu, ic
from tests.test_oracle_isolated import DT_TREE

pytestmark = pytest.mark.gpu

ECCS = np.array([0.05, 0.3, 0.6, 0.79, 0.8, 0.81, 0.9, 0.97])


def isolated_disk(n, seed):
    rng = np.random.default_rng(seed)
    ax = rng.uniform(0.6, 1.8, n)
    ecc = ECCS[np.arange(n) % len(ECCS)]
    inc = rng.uniform(0.0, 0.05, n)
    l, omg, OMG = (2 * np.pi * rng.random(n) for _ in range(3))
    pos, vel = ic.elements_to_posvel(ax, ecc, inc, l, omg, OMG)
    pp = np.zeros(n, dtype=abi.PARTICLE)
    pp["id"] = np.arange(n)
    pp["id_cluster"] = np.arange(n)
    pp["mass"] = 1e-12
    pp["pos"] = pos
    pp["vel"] = vel
    return pp


def run(oracle_lib, eps2_sun, n=512, seed=4):
    kw = dict(R_cut0=3.0, R_cut1=0.0, R_search0=1.1, R_search1=0.0, R_search2=1.1, R_search3=2.0,
              gamma=0.1, eta=0.01, eta_sun=0.01, eta_0=0.002, eta_sun0=0.002, dt_tree=DT_TREE,
              dt_min=2.0 ** -40, n_group_limit=64, eps2_sun=eps2_sun)
    pp = isolated_disk(n, seed)
    print("isolated particles: n=%d eps2_sun=%g" % (n, eps2_sun), flush=True)
    s = gpu.System(gpu.params(**kw), pp)
    s.initialize()
    pre = s.particles()
    assert np.all(pre["neighbor"]["number"] == 0)
    s.hard_part(True)
    post = s.particles()
    s.close()
    ref = pre.copy()
    oracle_lib.isolated(oracle_lib.params(**kw), ref, 0.0, DT_TREE)
    o = np.argsort(post["id"])
    r = np.argsort(ref["id"])
    return pre[np.argsort(pre["id"])], post[o], ref[r], oracle_lib.params(**kw)


FIELDS = ("pos", "vel", "acc_s", "jerk_s", "acc_d", "jerk_d", "phi_s", "phi_d", "time", "dt", "n_cluster")


def test_sun_only_hermite_bit_identical(oracle_lib):
    pre, got, ref, p = run(oracle_lib, eps2_sun=0.0)
    ecc = oracle_lib.eccentricity(p, pre)
    herm = ecc >= 0.8
    assert herm.sum() > 150 and (~herm).sum() > 200
    for f in FIELDS:
        assert np.array_equal(got[f][herm], ref[f][herm]), f
    assert np.all(got["time"] == DT_TREE)
    # Kepler branch: libm of device and host
    kep = ~herm
    for f in ("pos", "vel"):
        err = np.linalg.norm(got[f][kep] - ref[f][kep], axis=1) / np.linalg.norm(ref[f][kep], axis=1)
        assert err.max() < 1e-13, (f, err.max())
    assert np.all(got["time"][kep] == ref["time"][kep])
    assert np.mean(got["dt"][kep] == ref["dt"][kep]) > 0.99
    # kepler.h:23: the drift takes the fabs outcome; the int abs(int) one is far off at e = 0.79
    e79 = np.abs(ecc - 0.79) < 1e-6
    assert e79.sum() > 40
    alt = pre[e79].copy()
    oracle_lib.kepler_isolated(p, alt, 0.0, DT_TREE, int_abs=True)
    d_alt = np.linalg.norm(got["pos"][e79] - alt["pos"], axis=1)
    d_ref = np.linalg.norm(got["pos"][e79] - ref["pos"][e79], axis=1)
    assert d_ref.max() < 1e-12
    assert np.median(d_alt) > 1e-6


def test_softened_sun_all_hermite_bit_identical(oracle_lib):
    pre, got, ref, p = run(oracle_lib, eps2_sun=1e-6, n=256, seed=9)
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f
    assert np.all(got["time"] == DT_TREE)
