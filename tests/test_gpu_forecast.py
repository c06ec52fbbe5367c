"""Learned candidate-buffer sizing on the device system.

1. Results never depend on the forecast: a disk stepped with the network sizing the
   per-group candidate segments, with the uniform-rate ablation, and with the global
   buffer gives bit-identical particles and neighbour graphs.
2. What the trained network gains: on the same states, the network's per-group plan
   overflows far fewer i-groups (each overflow costs a candidate-only re-run of that
   group) than the ablation that gives every particle the previous mean count, at no
   more planned slots."""
import numpy as np
import pytest

import bench
from gplum_amd import gpu

pytestmark = pytest.mark.gpu

N = 60000


def disk_system(mode, n=N, seed=3):
    p = gpu.params(**bench.params_dict())
    s = gpu.System(p, bench.make_disk(n, seed))
    s.set_forecast(mode)
    return s


def by_id(pp):
    return pp[np.argsort(pp["id"], kind="stable")]


def test_results_identical_in_every_mode():
    runs = {}
    for mode in ("off", "network", "uniform"):
        s = disk_system(mode)
        s.initialize()
        s.reset_stats(timing=True)
        s.step(3)
        pp = s.particles()
        off, idx = s.neighbors()
        runs[mode] = (by_id(pp), gpu.neighbor_sets_by_id(pp, off, idx), s.stats())
        s.close()
    ref, nref, _ = runs["off"]
    for mode in ("network", "uniform"):
        pp, nset, st = runs[mode]
        print(mode, {k: st[k] for k in ("fc_calls", "fc_slots", "fc_ovf_groups", "fc_ovf_candidates",
                                        "fc_loss", "fc_ms")})
        assert st["fc_calls"] == 3
        assert np.array_equal(pp["id"], ref["id"])
        for f in ("pos", "vel", "acc", "acc_d", "dt", "time"):
            assert np.array_equal(pp[f], ref[f]), f
        assert nset == nref


def test_network_plan_beats_uniform_rate():
    """Train online for 6 steps, then compare the two plans over 4 more steps of the
    same trajectory (the plan does not change the trajectory)."""
    s = disk_system("network")
    s.initialize()
    s.step(6)
    theta, mu, isd, rounds = s.get_forecast()
    assert rounds >= 7
    state = s.particles()
    t_sys = s.stats()["time_sys"]
    s.close()
    res = {}
    for mode in ("network", "uniform"):
        q = gpu.System(gpu.params(**bench.params_dict()), state, time_sys=t_sys)
        q.set_forecast(mode)
        q.put_forecast(theta, mu, isd)
        q.initialize()
        q.reset_stats(timing=True)
        q.step(4)
        res[mode] = q.stats()
        q.close()
    nw, un = res["network"], res["uniform"]
    print("network: slots %d overflowed groups %d redone candidates %d loss %.4f"
          % (nw["fc_slots"], nw["fc_ovf_groups"], nw["fc_ovf_candidates"], nw["fc_loss"]))
    print("uniform: slots %d overflowed groups %d redone candidates %d"
          % (un["fc_slots"], un["fc_ovf_groups"], un["fc_ovf_candidates"]))
    assert nw["fc_slots"] <= 1.05 * un["fc_slots"]
    assert nw["fc_ovf_groups"] < 0.5 * un["fc_ovf_groups"]
