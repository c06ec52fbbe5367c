"""Pins the oracle's isolated-particle branch of HardSystem::timeIntegrate (hard.h:793-818)
against mathematics, since the reference ships no output for it:

* the Sun-only Hermite (orc_hermite_isolated, timeIntegrate_isolated, hermite.h:734-785)
  on orbits of e = 0.85 and 0.95 follows the exact Kepler solution (Danby's f and g
  functions, tests/test_oracle_hermite.py, independent of the oracle's Kepler solver)
  over a whole orbit, with the error falling about 2^4 when eta halves (4th order);
* with a softened Sun (eps2_sun != 0) it follows a high-order adaptive integration of
  the softened two-body problem (scipy DOP853), and conserves that problem's energy;
* the dispatch (hard.h:803-811) takes the Kepler drift only for e < 0.8 with an
  unsoftened Sun;
* solveKeplerEq's unqualified `abs` (kepler.h:23): with fabs the Newton refinement
  solves Kepler's equation to rounding at e = 0.79 (the largest eccentricity the Kepler
  drift sees); with int abs(int) it is skipped and the 6th-order series alone is off by
  up to 0.16 rad there.  The drift of the build (fabs) follows the exact orbit."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from gplum_amd import abi
from tests.test_oracle_hermite import kepler_fg, orbit_state

DT_TREE = 2.0 ** -4


def single(pos, vel):
    pp = np.zeros(1, dtype=abi.PARTICLE)
    pp["id"] = 7
    pp["mass"] = 1e-12
    pp["pos"] = pos
    pp["vel"] = vel
    pp["r_out"] = 1e-4
    pp["r_out_inv"] = 1e4
    pp["r_search"] = 1.1e-4
    return pp


def params(orc, eta, eps2_sun=0.0):
    return orc.params(eta=eta, eta_sun=eta, eta_0=eta / 5, eta_sun0=eta / 5, dt_tree=DT_TREE,
                      dt_min=2.0 ** -40, gamma=0.1, eps2_sun=eps2_sun)


def run_isolated_hermite(orc, p, pos, vel, nsteps):
    """Sun gravity and calcDeltatInitial at t = 0, then nsteps soft steps of the Sun-only
    Hermite; returns the particle and the time after every call."""
    pp = single(pos, vel)
    L = orc.lib()
    import ctypes
    L.orc_star_gravity(ctypes.byref(p), abi.ptr(pp))
    L.orc_calc_dt_initial(ctypes.byref(p), abi.ptr(pp))
    times = []
    for _ in range(nsteps):
        orc.hermite_isolated(p, pp)
        times.append(float(pp["time"][0]))
    return pp, times


@pytest.mark.parametrize("ecc", [0.85, 0.95])
def test_sun_only_hermite_follows_kepler_4th_order(oracle_lib, ecc):
    pos, vel = orbit_state(1.0, ecc, 0.05, 0.7)
    nsteps = 96                       # 6.0 time units: one orbit (2 pi) nearly, periapsis included
    t = nsteps * DT_TREE
    xr, vr = kepler_fg(pos, vel, t)
    errs = []
    for eta in (0.02, 0.01):
        pp, times = run_isolated_hermite(oracle_lib, params(oracle_lib, eta), pos, vel, nsteps)
        assert times == [DT_TREE * (k + 1) for k in range(nsteps)]
        dt = pp["dt"][0]
        assert dt > 0 and math.log2(dt) == round(math.log2(dt))
        assert math.fmod(pp["time"][0], dt) == 0.0
        err = np.linalg.norm(pp["pos"][0] - xr) / np.linalg.norm(xr)
        errv = np.linalg.norm(pp["vel"][0] - vr) / np.linalg.norm(vr)
        errs.append(err)
        assert err < 1e-6 and errv < 1e-6, (eta, err, errv)
        # the Sun terms are those of the final state (calcStarGravity after the loop)
        r = pp["pos"][0]
        rn = np.linalg.norm(r)
        np.testing.assert_allclose(pp["acc_s"][0], -r / rn ** 3, rtol=1e-14)
        assert pp["phi_s"][0] == pytest.approx(-1.0 / rn, rel=1e-15)
        assert pp["phi_d"][0] == 0.0
    assert 8.0 < errs[0] / errs[1] < 32.0, errs


def test_sun_only_hermite_softened_sun(oracle_lib):
    eps2 = 0.05 ** 2
    pos, vel = orbit_state(1.0, 0.85, 0.05, 0.7)      # periapsis 0.15: softening is ~10 %
    nsteps = 96
    t = nsteps * DT_TREE

    def rhs(_, y):
        x = y[:3]
        r2 = x @ x + eps2
        return np.concatenate([y[3:], -x / r2 ** 1.5])

    sol = solve_ivp(rhs, (0.0, t), np.concatenate([pos, vel]), method="DOP853", rtol=1e-13, atol=1e-15)
    xr = sol.y[:3, -1]
    unsoft = kepler_fg(pos, vel, t)[0]
    assert np.linalg.norm(unsoft - xr) > 1e-3        # the softening matters at this size

    def energy(x, v):
        return 0.5 * v @ v - 1.0 / math.sqrt(x @ x + eps2)

    e0 = energy(pos, vel)
    errs = []
    for eta in (0.02, 0.01):
        pp, _ = run_isolated_hermite(oracle_lib, params(oracle_lib, eta, eps2), pos, vel, nsteps)
        x, v = pp["pos"][0], pp["vel"][0]
        err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        de = abs(energy(x, v) - e0) / abs(e0)
        errs.append(err)
        assert err < 1e-6 and de < 1e-7, (eta, err, de)
        assert pp["phi_s"][0] == pytest.approx(-1.0 / math.sqrt(x @ x + eps2), rel=1e-15)
    assert 8.0 < errs[0] / errs[1] < 32.0, errs


def test_isolated_dispatch(oracle_lib):
    """hard.h:803-811: Kepler drift iff e < 0.8 and eps2_sun == 0, else Sun-only Hermite."""
    for ecc, eps2, kepler in ((0.79, 0.0, True), (0.81, 0.0, False), (0.3, 1e-8, False), (0.1, 0.0, True)):
        p = params(oracle_lib, 0.01, eps2)
        pos, vel = orbit_state(1.2, ecc, 0.03, 1.1)
        pp = single(pos, vel)
        pp["dt"] = DT_TREE / 64
        pp["time"] = 3 * DT_TREE
        assert (oracle_lib.eccentricity(p, pp)[0] < 0.8) == (ecc < 0.8)
        a, b = pp.copy(), pp.copy()
        oracle_lib.isolated(p, a, 3 * DT_TREE, 4 * DT_TREE)
        if kepler:
            oracle_lib.kepler_isolated(p, b, 3 * DT_TREE, 4 * DT_TREE)
        else:
            oracle_lib.hermite_isolated(p, b)
        b["n_cluster"] = 1
        assert a.tobytes() == b.tobytes(), (ecc, eps2)
        assert a["time"][0] == 4 * DT_TREE


def test_kepler_abs_quirk_at_e08(oracle_lib):
    """kepler.h:23 calls an unqualified abs on a double.  Both outcomes at e = 0.79: the
    fabs one (the build's) solves Kepler's equation to rounding; the int abs(int) one
    skips Newton and keeps the series' error, which is large at this eccentricity."""
    L = oracle_lib.lib()
    ecc = 0.79
    res_f, res_i = [], []
    for l in np.linspace(0.0, 2 * np.pi, 721):
        uf = L.orc_solve_kepler(l, ecc)
        ui = L.orc_solve_kepler_int_abs(l, ecc)
        res_f.append(abs(uf - ecc * math.sin(uf) - l))
        res_i.append(abs(ui - ecc * math.sin(ui) - l))
    assert max(res_f) < 4e-15
    assert max(res_i) > 1e-2
    # the drift itself: fabs follows the exact orbit, int abs does not
    p = params(oracle_lib, 0.01)
    pos, vel = orbit_state(1.0, ecc, 0.02, 2.4)
    xr, vr = kepler_fg(pos, vel, DT_TREE)
    a, b = single(pos, vel), single(pos, vel)
    oracle_lib.kepler_isolated(p, a, 0.0, DT_TREE)
    oracle_lib.kepler_isolated(p, b, 0.0, DT_TREE, int_abs=True)
    assert np.linalg.norm(a["pos"][0] - xr) < 1e-13
    assert np.linalg.norm(a["vel"][0] - vr) < 1e-12
    assert np.linalg.norm(b["pos"][0] - xr) > 1e-6
