"""GPU parity of the main-loop bookkeeping around the hot path (main_p3t.cpp:619-699)
against the oracle: removeParticlesOutOfBoundary (func.h:209-380) with its energy
bookkeeping and remove log, and the recompute every reset_step-th step.

Bars: the same particles removed in the same step (ids), identical neighbour graphs
every step, the remove log's ids, times and masses equal and positions/velocities within
1e-9 relative of the oracle's (the FP32 soft force sums differ in order), the energy
error (E - E0 - E_disp)/E0 below 1e-10 with the removed bodies carrying over 10 % of E0;
after a reset step the block steps equal those of an Initial correction of the same
state (bit for bit)."""
import numpy as np
import pytest

from gplum_amd import gpu, ic
from tests import helpers
from tests.test_oracle_loop import ring_with_leavers

pytestmark = pytest.mark.gpu


def parse_remove_log(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            f = line.split("\t")
            rows.append((float(f[0]), int(f[1]), [float(x) for x in f[2:]]))
    return rows


def test_boundary_removal_matches_oracle(oracle_lib, tmp_path):
    kw = helpers.CFG1
    pp = ring_with_leavers()
    log = str(tmp_path / "remove000001.dat")
    print("boundary removal: N=%d r_max 1.6 r_min 0.1" % len(pp), flush=True)
    s = gpu.System(gpu.params(**kw), pp)
    s.set_boundary(1.6, 0.1, 1024)
    s.set_event_logs(remove_path=log)
    s.initialize()
    run = oracle_lib.P3T(oracle_lib.params(**kw), pp, r_max=1.6, r_min=0.1, reset_step=1024)
    e0 = s.energy()
    for k in range(3):
        s.step(1)
        run.step()
        g = s.particles()
        assert len(g) == len(run.pp)
        off_g, idx_g = s.neighbors()
        off_o, idx_o = run.neighbor_graph()
        assert gpu.neighbor_sets_by_id(g, off_g, idx_g) == gpu.neighbor_sets_by_id(run.pp, off_o, idx_o)
        e = s.energy()
        err = (e[0] - e0[0] - e[5]) / e0[0]
        print("step %d: n=%d removed so far %d energy error %.3e" % (k, len(g), s.stats()["n_removed"], err))
        assert abs(err) < 1e-10
    st = s.stats()
    assert st["n_removed"] == 3 and st["n_recomputes"] >= 1
    assert abs(s.energy()[5]) > 0.1 * abs(e0[0])
    assert s.energy()[5] == pytest.approx(run.edisp.value, rel=1e-9)
    s.set_event_logs()            # closes the log
    rows = parse_remove_log(log)
    ref = run.removed[0]
    assert [r[1] for r in rows] == ref["id"].tolist() == [600, 601, 602]
    for (t, _, v), q in zip(rows, ref):
        assert t == pytest.approx(q["time"], abs=5e-9)           # printed with 8 decimals
        assert v[0] == q["mass"]
        # vector-relative: a component near zero carries the absolute error of the others
        assert np.linalg.norm(np.subtract(v[1:4], q["pos"])) <= 1e-9 * np.linalg.norm(q["pos"])
        assert np.linalg.norm(np.subtract(v[4:7], q["vel"])) <= 1e-9 * np.linalg.norm(q["vel"])
    s.close()
    run.close()


def test_reset_step_recompute_on_gpu(oracle_lib):
    kw = helpers.CFG1
    pp = ic.disk(3000, 1e-9, 0.98, 1.02, seed=12)
    s = gpu.System(gpu.params(**kw), pp)
    s.set_boundary(40.0, 0.1, 3)
    s.initialize()
    run = oracle_lib.P3T(oracle_lib.params(**kw), pp, reset_step=3)
    rec = 0
    for k in range(6):
        s.step(1)
        run.step()
        rec += run.loop.recomputed
        g = s.particles()
        off_g, idx_g = s.neighbors()
        off_o, idx_o = run.neighbor_graph()
        assert gpu.neighbor_sets_by_id(g, off_g, idx_g) == gpu.neighbor_sets_by_id(run.pp, off_o, idx_o)
    st = s.stats()
    assert st["n_recomputes"] == rec >= 2
    # the last step (istep 5) ended with the Initial correction: its block steps are those
    # of correctForceLongInitial on the same state
    g = s.particles()
    q = g.copy()
    po = oracle_lib.params(**kw)
    tree = oracle_lib.tree_build(po, q)
    oracle_lib.soft_force(po, q, tree)
    tree.free()
    oracle_lib.correct_force_long(po, q, initial=True)
    assert np.array_equal(q["dt"], g["dt"])
    s.close()
    run.close()
