"""The project's autotuner (gplum_system_autotune): it picks launch shapes among its
candidates and leaves the neighbour graph bit-exact; with the chosen shapes the FP32
force stays within the FP32 tolerance of the FP64 force on the same tree lists (the
i-group size changes which nodes are opened, as FDPS's n_group_limit does, so forces
of two group sizes differ by the tree approximation); the hard-part shapes change no
result at all."""
import numpy as np
import pytest

from gplum_amd import gpu, ic
from tests import helpers

pytestmark = pytest.mark.gpu

CFG1 = helpers.CFG1


def by_id(pp):
    return pp[np.argsort(pp["id"], kind="stable")]


def disk(n=20000, seed=5):
    return ic.disk(n, 1e-9, 0.9, 1.1, seed=seed)


def neighbour_sets(s):
    pp = s.particles()
    off, idx = s.neighbors()
    return gpu.neighbor_sets_by_id(pp, off, idx)


def test_autotune_choice_keeps_neighbours():
    p = gpu.params(collision=1, **CFG1)
    s = gpu.System(p, disk())
    s.initialize()
    # the autotuner recomputes the soft force as a step does (correctForceLong, not the
    # initial variant, whose re-search filter differs): compare with that
    s.soft_force(initial=False)
    before = neighbour_sets(s)
    t = s.autotune(trials=1)
    print("tuning:", t)
    assert t["n_group_limit"] in (64, 128, 256)
    assert t["soft_tile"] in (64, 128, 256)
    assert t["hard_wg_per_cu"] in (2, 4, 8, 16)
    assert t["hard_lds_max"] in (128, 256)
    assert t["soft_ms"] > 0.0
    assert s.get_tuning()["n_group_limit"] == t["n_group_limit"]
    # the soft force with the chosen shape: same neighbour graph; FP32 vs FP64 on the
    # same lists within the FP32 tolerance
    assert neighbour_sets(s) == before
    acc32 = by_id(s.particles())["acc"]
    s.soft_force(initial=False, use_f64=True)
    acc64 = by_id(s.particles())["acc"]
    assert helpers.rel_err_rms(acc32, acc64) < 1e-6
    s.step(2)
    s.close()


@pytest.mark.parametrize("groups,tile", [(64, 64), (64, 128), (128, 256), (256, 64)])
def test_soft_shapes_keep_neighbours(groups, tile):
    p = gpu.params(collision=1, **CFG1)
    pp = disk(seed=6)
    ref = gpu.System(p, pp)
    ref.initialize()
    alt = gpu.System(p, pp)
    alt.set_tuning(n_group_limit=groups, soft_tile=tile)
    alt.initialize()
    assert neighbour_sets(alt) == neighbour_sets(ref)
    a = by_id(alt.particles())["acc"]
    b = by_id(ref.particles())["acc"]
    if groups == CFG1["n_group_limit"]:
        # same groups, other tile: only the FP32 partial sums move
        assert helpers.rel_err_rms(a, b) < 1e-6
    alt.soft_force(initial=False, use_f64=False)
    a32 = by_id(alt.particles())["acc"]
    alt.soft_force(initial=False, use_f64=True)
    a64 = by_id(alt.particles())["acc"]
    assert helpers.rel_err_rms(a32, a64) < 1e-6
    ref.close()
    alt.close()


def test_hard_shapes_change_nothing():
    """Persistent grid and LDS binning threshold: bit-identical trajectories."""
    p = gpu.params(collision=1, **CFG1)
    pp = disk(seed=7)
    out = []
    for wg, lds in ((8, 256), (2, 128), (16, 128)):
        s = gpu.System(p, pp)
        s.set_tuning(hard_wg_per_cu=wg, hard_lds_max=lds)
        s.initialize()
        s.step(3)
        out.append(by_id(s.particles()))
        s.close()
    for q in out[1:]:
        for f in ("pos", "vel", "acc", "mass", "dt"):
            assert np.array_equal(q[f], out[0][f]), f
