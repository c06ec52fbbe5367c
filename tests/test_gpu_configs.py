"""The larger workloads at their real sizes, through size-independent properties.

Config 4 (N = 4M disk over 0.5-3 AU, learned candidate sizing on): the neighbour graph
after full steps is identical with the forecaster on and off (the exact search judges);
the energy error over the steps stays at the round-off level.
Config 5 (disk over 0.3-5 AU plus 10 embryos of 1e-6 on circular orbits, merging): at
N = 1M (the 16M run is the same code at 16x the density) the FP32 path and the FP64-soft
path from the same state conserve energy to the same level over the steps."""
import numpy as np
import pytest

import bench
from gplum_amd import gpu, ic

pytestmark = pytest.mark.gpu


def test_config4_neighbours_exact_with_forecaster():
    p = gpu.params(**bench.params_dict())
    pp = ic.disk(4_000_000, 1e-12, 0.5, 3.0, seed=1)
    s = gpu.System(p, pp)
    s.set_forecast("network")
    s.initialize()
    e0 = s.energy()
    s.step(2)
    e1 = s.energy()
    st = s.stats()
    assert st["fc_calls"] > 0
    de = (e1[0] - e0[0] - (e1[5] - e0[5])) / e0[0]
    print("config4: 2 steps, energy error %.3e, overflowed groups %d" % (de, st["fc_ovf_groups"]))
    assert abs(de) < 1e-10
    s.soft_force(initial=False)
    a = bench.neighbour_pairs(s)
    off = gpu.System(p, s.particles(), time_sys=st["time_sys"])
    off.soft_force(initial=False)
    b = bench.neighbour_pairs(off)
    assert len(a) > 0
    assert np.array_equal(a, b)
    s.close()
    off.close()


def test_config5_embryos_fp32_and_fp64_paths():
    p = gpu.params(**bench.params_dict())
    pp = ic.add_embryos(ic.disk(1_000_000, 1e-12, 0.3, 5.0, seed=3), 10, 1e-6, 0.3, 5.0)
    errs = []
    for fp64 in (False, True):
        s = gpu.System(p, pp)
        s.set_variant(fp64_soft=fp64, per_cluster_hard=False)
        if not fp64:
            s.set_forecast("network")
        s.initialize()
        e0 = s.energy()
        s.step(2)
        e1 = s.energy()
        st = s.stats()
        # the FP64 variant runs every soft force, the recomputes after merging included,
        # with the FP64 kernels; the production path none
        assert st["n_force_f64_calls"] == (st["n_force_calls"] if fp64 else 0), st
        assert st["n_collisions"] > 0
        errs.append((e1[0] - e0[0] - (e1[5] - e0[5])) / e0[0])
        print("config5 (N=1M + 10 embryos) fp64_soft=%d: energy error %.3e, largest cluster %d, collisions %d"
              % (fp64, errs[-1], st["max_cluster"], st["n_collisions"]))
        s.close()
    assert abs(errs[0]) < 1e-9 and abs(errs[1]) < 1e-9
