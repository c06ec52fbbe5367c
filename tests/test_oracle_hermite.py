"""Pins the oracle's block-timestep Hermite (orc_hermite_cluster, the restatement of
timeIntegrate_multi, hermite.h:224-479) against mathematics, since the reference ships
no Hermite output:

* test particles around the Sun follow the exact Kepler solution (Danby's f and g
  functions, written here independently of the oracle's own Kepler solver), and the
  error falls by about 2^4 when eta halves (4th order);
* a tight binary inside r_in feels the full Newtonian force: the total energy with the
  Sun is conserved to the integrator's accuracy;
* the centre of mass of a cluster, merging or not, follows the Kepler orbit of the
  total mass up to tidal terms (catches a wrong sign or a dropped mutual term, and a
  merge that does not conserve momentum); merged impactors ride with their target;
* block-step invariants: every dt is 2^-k, every time is a multiple of its dt."""
import numpy as np

from gplum_amd import abi
from tests import helpers

DT_TREE = 2.0 ** -4


def kepler_fg(r0, v0, t, mu=1.0):
    """Exact two-body propagation (elliptic) by the f and g functions."""
    r0 = np.asarray(r0, float)
    v0 = np.asarray(v0, float)
    rr = np.linalg.norm(r0)
    a = 1.0 / (2.0 / rr - v0 @ v0 / mu)
    n = np.sqrt(mu / a ** 3)
    ecosE = 1.0 - rr / a
    esinE = (r0 @ v0) / np.sqrt(mu * a)
    E0 = np.arctan2(esinE, ecosE)
    e = np.hypot(ecosE, esinE)
    M = E0 - esinE + n * t
    E = M
    for _ in range(60):
        E -= (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
    dE = E - E0
    f = 1.0 - a / rr * (1.0 - np.cos(dE))
    g = t - (dE - np.sin(dE)) / n
    r1 = f * r0 + g * v0
    r1n = np.linalg.norm(r1)
    fd = -np.sqrt(mu * a) * np.sin(dE) / (r1n * rr)
    gd = 1.0 - a / r1n * (1.0 - np.cos(dE))
    return r1, fd * r0 + gd * v0


def orbit_state(a, e, inc, phase):
    """Position/velocity at true anomaly `phase` of an orbit (mu = 1)."""
    p = a * (1 - e * e)
    r = p / (1 + e * np.cos(phase))
    pos = np.array([r * np.cos(phase), r * np.sin(phase), 0.0])
    vel = np.array([-np.sin(phase), e + np.cos(phase), 0.0]) * np.sqrt(1.0 / p)
    c, s = np.cos(inc), np.sin(inc)
    rot = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    return rot @ pos, rot @ vel


def cluster(pos, vel, mass, r_out, r_planet=None, f=1.0):
    n = len(mass)
    pp = np.zeros(n, dtype=abi.PARTICLE)
    pp["id"] = np.arange(n)
    pp["id_cluster"] = 0
    pp["mass"] = mass
    pp["pos"] = pos
    pp["vel"] = vel
    pp["r_out"] = r_out
    pp["r_out_inv"] = 1.0 / r_out
    pp["r_search"] = 1.1 * r_out
    pp["r_planet"] = 1e-9 if r_planet is None else r_planet
    pp["f"] = f
    ro = np.maximum.outer(pp["r_out"], pp["r_out"])
    w = ro ** 2 / pp["mass"][None, :]
    np.fill_diagonal(w, 0.0)
    pp["acc0"] = (n - 1) / w.sum(1)
    return pp


def integrate(oracle_lib, pp, eta, collision=0):
    p = oracle_lib.params(eta=eta, eta_sun=eta, eta_0=eta / 5, eta_sun0=eta / 5, dt_tree=DT_TREE,
                          dt_min=2.0 ** -40, gamma=0.1, collision=collision)
    out = pp.copy()
    oracle_lib.cluster_prepare(p, out, np.array([0, len(out)]))
    nc, ed, nb = oracle_lib.hermite_cluster(p, out)
    return out, nc, ed, nb


def check_block_steps(pp):
    k = -np.log2(pp["dt"])
    assert np.all(k == np.round(k)), pp["dt"]
    assert np.all(np.fmod(pp["time"], pp["dt"]) == 0.0)


def test_test_particles_follow_kepler_4th_order(oracle_lib):
    s1 = orbit_state(1.0, 0.1, 0.02, 0.3)
    s2 = orbit_state(1.4, 0.05, 0.01, 2.0)
    pp = cluster([s1[0], s2[0]], [s1[1], s2[1]], [1e-25, 1e-25], 1e-3)
    exact = [kepler_fg(*s, DT_TREE) for s in (s1, s2)]
    errs = []
    for eta in (0.02, 0.01):
        out, nc, ed, nb = integrate(oracle_lib, pp, eta)
        check_block_steps(out)
        assert np.all(out["time"] == DT_TREE)
        err = max(np.linalg.norm(out["pos"][i] - exact[i][0]) / np.linalg.norm(exact[i][0]) for i in range(2))
        errv = max(np.linalg.norm(out["vel"][i] - exact[i][1]) / np.linalg.norm(exact[i][1]) for i in range(2))
        print("eta %.3f block steps %d pos err %.3e vel err %.3e" % (eta, nb, err, errv))
        errs.append(err)
        assert err < 1e-8 and errv < 1e-8
    assert 8.0 < errs[0] / errs[1] < 40.0


def test_binary_inside_r_in_conserves_energy(oracle_lib):
    sc = orbit_state(1.0, 0.0, 0.0, 0.0)
    m, sep = 1e-8, 1e-4
    vrel = np.sqrt(2 * m / sep)
    d = np.array([sep / 2, 0.0, 0.0])
    dv = np.array([0.0, 0.0, vrel / 2])          # binary plane perpendicular to the orbit
    pp = cluster([sc[0] + d, sc[0] - d], [sc[1] + dv, sc[1] - dv], [m, m], 1e-2)

    def energy(q):
        r = np.linalg.norm(q["pos"], axis=1)
        r12 = np.linalg.norm(q["pos"][0] - q["pos"][1])
        return (0.5 * (q["mass"] * (q["vel"] ** 2).sum(1)).sum() - (q["mass"] / r).sum()
                - q["mass"][0] * q["mass"][1] / r12)

    e0 = energy(pp)
    e_bin = -m * m / (2 * sep)
    errs = []
    for eta in (0.02, 0.01):
        out, nc, ed, nb = integrate(oracle_lib, pp, eta)
        check_block_steps(out)
        de = abs(energy(out) - e0) / abs(e_bin)
        print("eta %.3f block steps %d |dE|/|E_bin| %.3e" % (eta, nb, de))
        errs.append(de)
        assert de < 1e-4
    assert errs[1] < errs[0]


def com_check(oracle_lib, pp, collision):
    m = pp["mass"]
    M = m.sum()
    rc = (m[:, None] * pp["pos"]).sum(0) / M
    vc = (m[:, None] * pp["vel"]).sum(0) / M
    out, nc, ed, nb = integrate(oracle_lib, pp, 0.01, collision=collision)
    alive = (out["flags"] & 1) == 0
    check_block_steps(out[alive])
    # a merged body keeps its impactors as dead entries with the target's id, position
    # and velocity until MergeParticle (func.h:160-207) adds their masses after the step
    for d in np.nonzero(~alive)[0]:
        t = np.nonzero(alive & (out["id"] == out["id"][d]))[0]
        assert len(t) == 1
        assert np.array_equal(out["pos"][d], out["pos"][t[0]])
        assert np.array_equal(out["vel"][d], out["vel"][t[0]])
    assert np.array_equal(out["mass"], m)
    rc1 = (m[:, None] * out["pos"]).sum(0) / M
    vc1 = (m[:, None] * out["vel"]).sum(0) / M
    r_ex, v_ex = kepler_fg(rc, vc, DT_TREE)
    er = np.linalg.norm(rc1 - r_ex) / np.linalg.norm(r_ex)
    ev = np.linalg.norm(vc1 - v_ex) / np.linalg.norm(v_ex)
    return nc, ed, er, ev


def test_cluster_centre_of_mass_follows_kepler(oracle_lib):
    rng = np.random.default_rng(4)
    sc = orbit_state(1.0, 0.02, 0.01, 1.0)
    n = 12
    pos = sc[0] + rng.normal(0, 2e-4, (n, 3))
    vel = sc[1] + rng.normal(0, 1e-4, (n, 3))
    pp = cluster(pos, vel, 10 ** rng.uniform(-10, -8, n), 3e-3)
    nc, ed, er, ev = com_check(oracle_lib, pp, collision=0)
    print("no merging: COM pos err %.3e vel err %.3e" % (er, ev))
    assert nc == 0
    assert er < 1e-7 and ev < 1e-7


def test_merging_conserves_momentum(oracle_lib):
    rng = np.random.default_rng(8)
    sc = orbit_state(1.0, 0.02, 0.01, 1.0)
    n = 10
    pos = sc[0] + rng.normal(0, 1e-4, (n, 3))
    vel = sc[1] + rng.normal(0, 3e-4, (n, 3))
    m = 10 ** rng.uniform(-10, -8, n)
    pp = cluster(pos, vel, m, 1e-2, r_planet=5e-6, f=10.0)
    nc, ed, er, ev = com_check(oracle_lib, pp, collision=1)
    print("merging: %d collisions, edisp %.3e, COM pos err %.3e vel err %.3e" % (nc, ed, er, ev))
    assert nc >= 1
    assert er < 1e-7 and ev < 1e-7


def body_energy(q):
    """Kinetic + Sun + full pairwise potential, entries with one id merged into one body."""
    ids = np.unique(q["id"])
    m = np.array([q["mass"][q["id"] == i].sum() for i in ids])
    first = [np.nonzero(q["id"] == i)[0][0] for i in ids]
    x, v = q["pos"][first], q["vel"][first]
    e = 0.5 * (m * (v ** 2).sum(1)).sum() - (m / np.linalg.norm(x, axis=1)).sum()
    for a in range(len(m)):
        for b in range(a + 1, len(m)):
            e -= m[a] * m[b] / np.linalg.norm(x[a] - x[b])
    return e


def test_merging_energy_change_equals_edisp(oracle_lib):
    """Two bodies inside r_in = gamma r_out (full Newton force) collide head-on: the
    full energy changes exactly by the dissipation the oracle reports
    (calcEnergyDissipation, collisionA.h:358-669).  Two bodies only: with perfect merging
    the reference skips the change of the interaction with third bodies (its loop runs
    to id_c_frag, which is 0 without fragments, collisionA.h:562), so the identity holds
    only without them."""
    sc = orbit_state(1.0, 0.0, 0.0, 0.0)
    d = np.array([1e-4, 0.0, 0.0])
    dv = np.array([-3e-3, 0.0, 0.0])
    pp = cluster([sc[0] + d, sc[0] - d], [sc[1] + dv, sc[1] - dv], np.array([1e-9, 2e-9]), 1e-2,
                 r_planet=1e-6, f=10.0)
    e0 = body_energy(pp)
    out, nc, ed, nb = integrate(oracle_lib, pp, 0.01, collision=1)
    e1 = body_energy(out)
    print("collisions %d edisp %.6e  E1-E0 %.6e" % (nc, ed, e1 - e0))
    assert nc == 1 and ed < 0.0
    assert abs((e1 - e0) - ed) < 1e-6 * abs(ed)


def test_collision_record_of_a_head_on_merger(oracle_lib):
    """The collision log's record (Collision0::write2File, collisionA.h:122-148) of the
    head-on pair above: the lighter body is the impactor, the masses are the bodies',
    the time lies on the block-step grid inside the soft step, the two bodies are closer
    than f (R_imp + R_tar) (the merging criterion, hermite.h:98-196), and the centre of
    mass does not move in the merger (the relative shifts, zero in exact arithmetic, are
    round-off)."""
    sc = orbit_state(1.0, 0.0, 0.0, 0.0)
    d = np.array([1e-4, 0.0, 0.0])
    dv = np.array([-3e-3, 0.0, 0.0])
    pp = cluster([sc[0] + d, sc[0] - d], [sc[1] + dv, sc[1] - dv], np.array([1e-9, 2e-9]), 1e-2,
                 r_planet=1e-6, f=10.0)
    p = oracle_lib.params(eta=0.01, eta_sun=0.01, eta_0=0.002, eta_sun0=0.002, dt_tree=DT_TREE,
                          dt_min=2.0 ** -40, gamma=0.1, collision=1)
    out = pp.copy()
    oracle_lib.cluster_prepare(p, out, np.array([0, 2]))
    nc, ed, nb, rec = oracle_lib.hermite_cluster_log(p, out)
    assert nc == 1 and len(rec) == 1
    r = dict(zip(oracle_lib.COLREC_FIELDS, rec[0]))
    print("merger at t = %.10e, |dx| %.3e, shifts %.2e %.2e" % (
        r["time"], np.linalg.norm([r["pos_imp_x"] - r["pos_tar_x"], r["pos_imp_y"] - r["pos_tar_y"],
                                    r["pos_imp_z"] - r["pos_tar_z"]]), r["dpos_g"], r["dvel_g"]))
    assert (r["id_imp"], r["id_tar"]) == (0, 1)
    assert (r["mass_imp"], r["mass_tar"]) == (1e-9, 2e-9)
    assert 0.0 < r["time"] < DT_TREE
    assert float(np.ldexp(r["time"], 40)).is_integer()
    dx = np.array([r["pos_imp_" + c] - r["pos_tar_" + c] for c in "xyz"])
    assert np.linalg.norm(dx) < 10.0 * (1e-6 + 1e-6)
    assert r["dpos_g"] < 4e-16 and r["dvel_g"] < 4e-16
    # the same merger as the plain integration (same state afterwards)
    out2 = pp.copy()
    oracle_lib.cluster_prepare(p, out2, np.array([0, 2]))
    nc2, ed2, nb2 = oracle_lib.hermite_cluster(p, out2)
    assert (nc2, ed2, nb2) == (nc, ed, nb) and np.array_equal(out2["pos"], out["pos"])


def test_config1_run_energy_drift(oracle_lib, init3000):
    """The reference's sample disk (INIT3000, 3000 planetesimals, monopole, merging) over
    64 soft steps of 2^-4: |dE| = |(E - E0 - E_disp) / E0| (energy.h:65-67) stays below
    1e-9 at every 8th step.  This is the oracle's own drift, the bar the device system
    is held to."""
    hdr, pp = init3000
    p = oracle_lib.params(quadrupole=0, collision=1, **helpers.CFG1)
    run = oracle_lib.P3T(p, pp)
    e0 = oracle_lib.energy(run.pp)[0]
    worst = 0.0
    for k in range(64):
        run.step()
        if k % 8 == 7:
            de = abs((oracle_lib.energy(run.pp)[0] - e0 - run.edisp.value) / e0)
            worst = max(worst, de)
    run.close()
    print("max |dE| over 64 steps %.3e" % worst)
    assert worst < 1e-9


def test_cube_root_within_one_ulp(oracle_lib):
    """The merged-radius cube root (shared with the kernels) is within 1 ulp of the
    true root over the range mass ratios take, and exact on perfect cubes."""
    rng = np.random.default_rng(0)
    x = np.concatenate([10 ** rng.uniform(-6, 6, 20000), [1.0, 8.0, 27.0, 0.125, 2.0, 1e-300, 1e300]])
    got = np.array([oracle_lib.lib().orc_cbrt_exact_ops(float(v)) for v in x])
    ref = np.cbrt(x)
    ulp = np.spacing(ref)
    assert np.all(np.abs(got - ref) <= ulp)
    assert got[-7] == 1.0 and got[-6] == 2.0 and got[-5] == 3.0 and got[-4] == 0.5


def test_exp_within_one_ulp(oracle_lib):
    """exp used by the velocity-dispersion smoothing (shared with the kernels) is within
    1 ulp of libm's over the arguments it takes (-(ddr/dr)^2 <= 0), exact at 0, and 0
    below the subnormal range."""
    rng = np.random.default_rng(1)
    x = np.concatenate([-rng.uniform(0, 1, 20000), -rng.uniform(0, 740, 20000), [0.0, -1e-300, -800.0]])
    got = np.array([oracle_lib.lib().orc_exp_exact_ops(float(v)) for v in x])
    ref = np.exp(x)
    assert np.all(np.abs(got - ref) <= np.spacing(ref))
    assert got[-3] == 1.0 and got[-2] == 1.0 and got[-1] == 0.0
