"""GPU parity of the batched FP64 Hermite (persistent and per-cluster launches) against the
oracle's timeIntegrate_multi restatement, cluster by cluster.

Both sides use the reference's operation order without fused multiply-add, so the whole
state (block steps, times, positions, velocities, forces, potentials) must agree bit for
bit; collisions (perfect merging) must happen in the same block step between the same
bodies, with the same merged radii and dissipated energy."""
import numpy as np
import pytest

from gplum_amd import gpu, ic

pytestmark = pytest.mark.gpu

# dt_min = 2^-20: without softening or collisions a close encounter would otherwise
# drive the block step down to 2^-30 (millions of block steps for one cluster)
HARD = dict(eta=0.01, eta_sun=0.01, eta_0=0.002, eta_sun0=0.002, dt_tree=2.0 ** -4,
            dt_min=2.0 ** -20, gamma=0.1)


def make_clusters(oracle_lib, n_clusters, seed, size_max, collision, f=1.0, m_hi=1e-9, planar=False):
    po = oracle_lib.params(collision=collision, **HARD)
    pg = gpu.params(collision=collision, **HARD)
    pp, off = ic.clusters(n_clusters, seed=seed, size_max=size_max, m_hi=m_hi)
    pp["f"] = f
    if planar:
        pp["pos"][:, 2] = 0.0     # z and v_z exactly zero: signed zeros reach the results
    # acc0 as correctForceLong forms it: number / sum r_out^2 / m_j over the cluster
    for c in range(n_clusters):
        a, b = off[c], off[c + 1]
        ro = np.maximum.outer(pp["r_out"][a:b], pp["r_out"][a:b])
        w = ro ** 2 / pp["mass"][a:b][None, :]
        np.fill_diagonal(w, 0.0)
        pp["acc0"][a:b] = (b - a - 1) / w.sum(1)
    oracle_lib.cluster_prepare(po, pp, off)
    return po, pg, pp, off


@pytest.fixture(scope="module")
def no_collision_case(oracle_lib):
    """60 clusters (2..300 bodies, one above the 256-body LDS limit) and the oracle result."""
    po, pg, pp, off = make_clusters(oracle_lib, 60, seed=2, size_max=300, collision=0)
    assert np.diff(off).max() > 256
    ref = pp.copy()
    nc, ed, nb_ref = oracle_lib.hermite_clusters(po, ref, off)
    return pg, pp, off, ref, nb_ref


@pytest.mark.parametrize("persistent", [True, False])
def test_batched_hermite_matches_oracle(no_collision_case, persistent):
    pg, pp, off, ref, nb_ref = no_collision_case
    got = pp.copy()
    ed_g, nb_g, ms = gpu.hard_integrate(pg, got, off, persistent=persistent)
    print("clusters %d particles %d block steps %d (oracle %d) kernel %.2f ms"
          % (len(off) - 1, len(pp), nb_g, nb_ref, ms))
    assert nb_g == nb_ref
    assert np.array_equal(got["time"], ref["time"])
    assert np.array_equal(got["dt"], ref["dt"])
    rel = np.abs(got["pos"] - ref["pos"]).max() / np.abs(ref["pos"]).max()
    relv = np.abs(got["vel"] - ref["vel"]).max() / np.abs(ref["vel"]).max()
    print("max rel diff pos %.3e vel %.3e" % (rel, relv))
    # same operations in the same order on both sides: bit for bit (phi_d goes through
    # the library log of the changeover potential, which may differ in the last bit)
    for f in ("pos", "vel", "acc_d", "jerk_d", "acc_s", "jerk_s", "phi_s"):
        assert np.array_equal(got[f], ref[f]), f
    assert np.allclose(got["phi_d"], ref["phi_d"], rtol=1e-13, atol=0.0)


def test_hermite_perfect_merging_matches_oracle(oracle_lib):
    # inflated radii (f) make bodies collide within the soft step
    po, pg, pp, off = make_clusters(oracle_lib, 40, seed=9, size_max=24, collision=1, f=300.0,
                                    m_hi=1e-8)
    ref = pp.copy()
    nc, ed, nb_ref = oracle_lib.hermite_clusters(po, ref, off)
    got = pp.copy()
    ed_g, nb_g, ms = gpu.hard_integrate(pg, got, off, persistent=True)
    dead_ref = (ref["flags"] & 1) != 0
    dead_got = (got["flags"] & 1) != 0
    print("collisions (oracle) %d, dead %d, edisp %.3e vs %.3e" % (nc, dead_ref.sum(), ed, ed_g))
    # cluster by cluster, so that a mismatch names its cluster
    bad = []
    for c in range(len(off) - 1):
        a, b = int(off[c]), int(off[c + 1])
        r = pp[a:b].copy()
        nc_c, ed_c, nb_c = oracle_lib.hermite_cluster(po, r)
        g = pp[a:b].copy()
        ed_gc, nb_gc, _ = gpu.hard_integrate(pg, g, np.array([0, b - a]), persistent=True)
        dx = np.abs(g["pos"] - r["pos"]).max() / np.abs(r["pos"]).max()
        if nb_gc != nb_c or dx > 0.:
            bad.append((c, b - a, nc_c, nb_c, nb_gc, float(dx), float(np.abs(g["dt"] - r["dt"]).max())))
    print("clusters differing (c, n, collisions, steps oracle, steps gpu, pos diff, dt diff):", bad)
    assert nc > 0
    assert np.array_equal(dead_got, dead_ref)
    assert np.array_equal(got["id"], ref["id"])
    assert not bad
    assert nb_g == nb_ref
    assert ed_g == ed
    for f in ("pos", "vel", "acc_d", "jerk_d", "mass", "r_planet", "dt", "time"):
        assert np.array_equal(got[f], ref[f]), f


def test_repeated_ids_match_oracle(oracle_lib):
    """A cluster member that repeats another member's id is skipped as a partner (the
    oracle's same-id test); such clusters leave the register path for the global-memory
    path, which keeps the test."""
    po, pg, pp, off = make_clusters(oracle_lib, 30, seed=17, size_max=40, collision=0)
    rep = 0
    for c in range(len(off) - 1):
        if off[c + 1] - off[c] >= 3 and c % 2 == 0:
            pp["id"][off[c] + 2] = pp["id"][off[c]]
            rep += 1
    assert rep > 3
    ref = pp.copy()
    nc, ed, nb_ref = oracle_lib.hermite_clusters(po, ref, off)
    got = pp.copy()
    ed_g, nb_g, ms = gpu.hard_integrate(pg, got, off, persistent=True)
    print("clusters with a repeated id %d, block steps %d (oracle %d)" % (rep, nb_g, nb_ref))
    assert nb_g == nb_ref
    for f in ("pos", "vel", "acc_d", "jerk_d", "acc_s", "jerk_s", "phi_s", "dt", "time"):
        assert np.array_equal(got[f], ref[f]), f


def test_fast_time_step_criterion_matches_exact():
    """The Hermite time-step rule with the hardware-sqrt criterion and the exact fallback
    near the step grid gives the correctly rounded rule's step, bit for bit, on 4M random
    states (accelerations and derivatives over many decades, steps dt_tree 2^-k)."""
    pg = gpu.params(collision=0, **HARD)
    mism, fb = gpu.hermite_dt_selfcheck(pg, 4_000_000, seed=7)
    print("mismatches %d, fallbacks to the exact criterion %d of 4M" % (mism, fb))
    assert mism == 0
    assert fb > 0      # the exact branch ran


def _same_state(a, b, fields=("pos", "vel", "acc_d", "jerk_d", "acc_s", "jerk_s", "phi_d", "phi_s", "dt", "time",
                              "mass", "r_planet", "id", "flags")):
    return [f for f in fields if not np.array_equal(a[f], b[f])]


def test_graph_replay_matches_eager(no_collision_case):
    """The persistent launch captured into a hipGraph (state restore, queue reset by a
    kernel, size-class kernels on two streams): every replay reproduces the first bit
    for bit (checked inside the call with synchronisations between replays), and the
    first equals the eager launch."""
    pg, pp, off, ref, nb_ref = no_collision_case
    eager = pp.copy()
    ed_e, nb_e, ms_e = gpu.hard_integrate(pg, eager, off, persistent=True)
    got = pp.copy()
    ed_g, nb_g, ms_g, bad = gpu.hard_integrate_graph(pg, got, off, replays=3)
    print("eager %.2f ms, graph replay %.2f ms, mismatching replays %d" % (ms_e, ms_g, bad))
    assert bad == 0
    assert nb_g == nb_e == nb_ref
    assert _same_state(got, eager) == []


def test_graph_replay_with_merging(oracle_lib):
    """Replays restore ids, flags and radii changed by perfect merging."""
    po, pg, pp, off = make_clusters(oracle_lib, 24, seed=9, size_max=24, collision=1, f=300.0, m_hi=1e-8)
    eager = pp.copy()
    ed_e, nb_e, _ = gpu.hard_integrate(pg, eager, off, persistent=True)
    assert ((eager["flags"] & 1) != 0).any()
    got = pp.copy()
    ed_g, nb_g, _, bad = gpu.hard_integrate_graph(pg, got, off, replays=3)
    assert bad == 0
    assert nb_g == nb_e and ed_g == ed_e
    assert _same_state(got, eager) == []


@pytest.mark.parametrize("planar", [False, True])
def test_small_clusters_lane_per_term(oracle_lib, planar):
    """Clusters of 2 to 4 bodies (one lane per particle, term and component): bit-identical
    to the oracle and to the one-lane-per-particle path; in the planar case z and v_z stay
    exactly zero and the Sun's z-terms are -0, so the sign handling of the quad dot
    products shows in the compared bits."""
    po, pg, pp, off = make_clusters(oracle_lib, 300, seed=23, size_max=4, collision=0, planar=planar)
    assert np.diff(off).min() == 2 and np.diff(off).max() == 4
    ref = pp.copy()
    nc, ed, nb_ref = oracle_lib.hermite_clusters(po, ref, off)
    got = pp.copy()
    ed_g, nb_g, ms = gpu.hard_integrate(pg, got, off, persistent=True)
    lpp = pp.copy()
    ed_l, nb_l, ms_l = gpu.hard_integrate(pg, lpp, off, persistent=True, lane_per_particle=True)
    print("planar %s: %d clusters, block steps %d (oracle %d), kernel %.2f ms (one lane per particle %.2f ms)"
          % (planar, len(off) - 1, nb_g, nb_ref, ms, ms_l))
    assert nb_g == nb_ref == nb_l
    assert _same_state(got, ref, fields=("pos", "vel", "acc_d", "jerk_d", "acc_s", "jerk_s", "phi_s", "dt",
                                         "time")) == []
    assert np.allclose(got["phi_d"], ref["phi_d"], rtol=1e-13, atol=0.0)
    assert _same_state(got, lpp) == []
    if planar:
        assert (got["pos"][:, 2] == 0.0).all() and np.signbit(got["acc_s"][:, 2]).any()


def test_small_clusters_lane_per_term_collisions(oracle_lib):
    """Clusters of 2 to 4 bodies with inflated radii: the block step of a collision hands
    the cluster to the global-memory path; the same merges as the oracle."""
    po, pg, pp, off = make_clusters(oracle_lib, 200, seed=29, size_max=4, collision=1, f=300.0, m_hi=1e-8)
    ref = pp.copy()
    nc, ed, nb_ref = oracle_lib.hermite_clusters(po, ref, off)
    got = pp.copy()
    ed_g, nb_g, ms = gpu.hard_integrate(pg, got, off, persistent=True)
    print("collisions %d, block steps %d (oracle %d)" % (nc, nb_g, nb_ref))
    assert nc > 10
    assert nb_g == nb_ref
    assert ed_g == ed
    assert _same_state(got, ref, fields=("pos", "vel", "acc_d", "jerk_d", "mass", "r_planet", "dt", "time", "id",
                                         "flags")) == []


def test_giant_path_matches_oracle(no_collision_case):
    """The whole-GPU path for giant clusters (here forced for every cluster above 64
    particles): same operations, same orders, so bit-identical to the oracle."""
    pg, pp, off, ref, nb_ref = no_collision_case
    got = pp.copy()
    ed_g, nb_g, ms = gpu.hard_integrate(pg, got, off, persistent=True, giant=True)
    print("giant path: %d block steps in %.1f ms" % (nb_g, ms))
    assert nb_g == nb_ref
    for f in ("pos", "vel", "acc_d", "jerk_d", "acc_s", "jerk_s", "phi_s", "dt", "time"):
        assert np.array_equal(got[f], ref[f]), f
    assert np.allclose(got["phi_d"], ref["phi_d"], rtol=1e-13, atol=0.0)


def test_giant_path_with_merging(oracle_lib):
    po, pg, pp, off = make_clusters(oracle_lib, 6, seed=11, size_max=120, collision=1, f=100.0, m_hi=1e-8)
    a = pp.copy()
    ed_a, nb_a, _ = gpu.hard_integrate(pg, a, off, persistent=True)
    b = pp.copy()
    ed_b, nb_b, _ = gpu.hard_integrate(pg, b, off, persistent=True, giant=True)
    print("merging: %d collisions-worth of dissipation %.3e, block steps %d" % (((a["flags"] & 1) != 0).sum(), ed_a, nb_a))
    assert nb_a == nb_b and ed_a == ed_b
    assert _same_state(a, b) == []
