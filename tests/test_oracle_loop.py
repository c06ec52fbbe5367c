"""The oracle's main-loop bookkeeping around the hot path (main_p3t.cpp:619-699):
removeParticlesOutOfBoundary (func.h:209-380) and the recompute after a collision, a
removal or every reset_step-th step.

Pinned by the energy identity the reference's energy log relies on: the removed particles'
energy (kinetic, Sun, soft and hard potential, minus the mutual potential of removed pairs
counted twice) is booked as dissipated, so (E - E0 - E_disp) / E0 stays at the
integration error while E itself jumps by far more.  The removed set is the particles
outside [r_min, r_max]."""
import numpy as np

from gplum_amd import abi, ic
from tests import helpers


def ring_with_leavers(seed=3):
    """A 600-body ring at 1 AU plus a heavy bound pair beyond 1.6 AU (neighbours of each
    other, so their mutual potential sits in phi_d) and one body inside 0.1 AU."""
    pp = ic.disk(600, 1e-9, 0.95, 1.05, seed=seed)
    extra = np.zeros(3, dtype=abi.PARTICLE)
    extra["id"] = [600, 601, 602]
    extra["id_cluster"] = extra["id"]
    extra["mass"] = [2e-7, 2e-7, 1e-8]
    r1 = 1.62
    v1 = np.sqrt(1.0 / r1)
    d, dv = 3e-4, np.sqrt(4e-7 / 3e-4) / 2
    extra["pos"] = [[r1 + d / 2, 0.0, 0.0], [r1 - d / 2, 0.0, 0.0], [0.0, 0.08, 0.0]]
    extra["vel"] = [[0.0, v1 + dv, 0.0], [0.0, v1 - dv, 0.0], [-np.sqrt(1.0 / 0.08), 0.0, 0.0]]
    return np.concatenate([pp, extra])


def test_boundary_removal_keeps_energy_error(oracle_lib):
    p = oracle_lib.params(**helpers.CFG1)
    pp = ring_with_leavers()
    run = oracle_lib.P3T(p, pp, r_max=1.6, r_min=0.1, reset_step=1024)
    e0 = oracle_lib.energy(run.pp)
    errs = []
    for k in range(3):
        run.step()
        e = oracle_lib.energy(run.pp)
        assert np.allclose(e, np.array(run.loop.e_now), rtol=0, atol=0)   # energy at the end of the step
        errs.append((e[0] - e0[0] - run.edisp.value) / e0[0])
        if k == 0:
            gone = run.removed[0]
            assert run.loop.n_remove == 3 and run.loop.recomputed == 1
            assert gone["id"].tolist() == [600, 601, 602]
            assert len(run.pp) == 600
            assert set(run.pp["id"].tolist()) == set(range(600))
            # the removed bodies carried a large part of the energy
            assert abs(run.edisp.value) > 0.1 * abs(e0[0])
        else:
            assert run.loop.n_remove == 0
    print("energy errors", errs)
    assert max(abs(x) for x in errs) < 1e-10


def test_removed_pair_term(oracle_lib):
    """The mutual potential of the removed pair is booked once: the bookkeeping with the
    pair term dropped would leave an error of m_i m_j / (r E0) ~ 1e-4."""
    p = oracle_lib.params(**helpers.CFG1)
    pp = ring_with_leavers()
    run = oracle_lib.P3T(p, pp, r_max=1.6, r_min=0.1)
    e0 = oracle_lib.energy(run.pp)[0]
    run.step()
    gone = run.removed[0]
    i, j = (np.nonzero(gone["id"] == k)[0][0] for k in (600, 601))
    dr = gone["pos"][i] - gone["pos"][j]
    pair = gone["mass"][i] * gone["mass"][j] / np.sqrt(dr @ dr)
    e1 = oracle_lib.energy(run.pp)[0]
    err = (e1 - e0 - run.edisp.value) / e0
    assert abs(err) < 1e-10
    assert abs(pair / e0) > 1e-5


def test_reset_step_recompute(oracle_lib):
    """Every reset_step-th step (istep % reset_step == reset_step - 1) ends with setCutoffRadii
    and the Initial force correction, which re-initialises every block step."""
    p = oracle_lib.params(**helpers.CFG1)
    pp = ic.disk(800, 1e-9, 0.95, 1.05, seed=8)
    run = oracle_lib.P3T(p, pp, reset_step=3)
    flags = []
    e0 = oracle_lib.energy(run.pp)[0]
    for _ in range(6):
        before = run.loop.istep
        ncol = run.step()
        flags.append((before, run.loop.recomputed, ncol))
    assert [f[0] for f in flags] == list(range(6))
    for istep, rec, ncol in flags:
        assert rec == (1 if (istep % 3 == 2 or ncol) else 0)
    # the recompute leaves the state of an Initial correction: the block steps equal those
    # calcDeltatInitial gives from the state, recomputed independently
    q = run.pp.copy()
    oracle_lib.set_cutoff_radii(p, q)
    tree = oracle_lib.tree_build(p, q)
    oracle_lib.soft_force(p, q, tree)
    tree.free()
    oracle_lib.correct_force_long(p, q, initial=True)
    assert np.array_equal(q["dt"], run.pp["dt"])
    e1 = oracle_lib.energy(run.pp)[0]
    assert abs((e1 - e0 - run.edisp.value) / e0) < 1e-10
