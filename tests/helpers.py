"""Shared builders for the tests: reference-layout EPI/EPJ/SPJ arrays from particles."""
import numpy as np

from gplum_amd import abi, ic


def to_epi(pp, idx):
    e = np.zeros(len(idx), dtype=abi.EPI)
    e["id_local"] = idx
    e["pos"] = pp["pos"][idx]
    e["r_out"] = pp["r_out"][idx]
    e["r_search"] = pp["r_search"][idx]
    return e


def to_epj(pp, idx):
    e = np.zeros(len(idx), dtype=abi.EPJ)
    e["id_local"] = idx
    e["pos"] = pp["pos"][idx]
    e["r_out"] = pp["r_out"][idx]
    e["r_search"] = pp["r_search"][idx]
    e["id"] = pp["id"][idx]
    e["mass"] = pp["mass"][idx]
    e["vel"] = pp["vel"][idx]
    e["acc_d"] = pp["acc_d"][idx]
    return e


def disk_with_radii(orc, n, seed=1, a_in=0.5, a_out=2.0, mass=1e-12, **pkw):
    """Config-2 disk law at size n, radii from the oracle's setCutoffRadii."""
    pp = ic.disk(n, mass, a_in, a_out, seed=seed)
    p = orc.params(**pkw)
    orc.set_cutoff_radii(p, pp)
    return p, pp


def group_lists(orc, p, pp, tree, g):
    """EPI, EPJ, SPJ(quad), SPJ(mono) arrays of i-group g of an oracle tree."""
    groups, epo, epl, spo, spn, perm = tree.lists()
    node, start, cnt = groups[g]
    iidx = perm[start:start + cnt]
    ns = tree.arr("node_start", tree.n_node)
    nc = tree.arr("node_count", tree.n_node)
    jidx = np.concatenate([perm[ns[c]:ns[c] + nc[c]] for c in epl[epo[g]:epo[g + 1]]]) \
        if epo[g + 1] > epo[g] else np.zeros(0, np.int64)
    sp = spn[spo[g]:spo[g + 1]]
    mass = tree.arr("node_mass", tree.n_node).astype(np.float64) if False else \
        np.ctypeslib.as_array(tree.node_mass, shape=(tree.n_node,)).copy()
    com = np.ctypeslib.as_array(tree.node_com, shape=(3 * tree.n_node,)).reshape(-1, 3).copy()
    quad = np.ctypeslib.as_array(tree.node_quad, shape=(6 * tree.n_node,)).reshape(-1, 6).copy()
    spq = np.zeros(len(sp), dtype=abi.SPJ_QUAD)
    spq["mass"] = mass[sp]
    spq["pos"] = com[sp]
    spq["quad"] = quad[sp]
    spm = np.zeros(len(sp), dtype=abi.SPJ_MONO)
    spm["mass"] = mass[sp]
    spm["pos"] = com[sp]
    return to_epi(pp, iidx), to_epj(pp, jidx), spq, spm


def rel_err_rms(a, ref):
    """rms over particles of |a - ref| / |ref| (vectors: row norms)."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    if a.ndim == 2:
        num = np.linalg.norm(a - ref, axis=1)
        den = np.linalg.norm(ref, axis=1)
    else:
        num = np.abs(a - ref)
        den = np.abs(ref)
    return float(np.sqrt(np.mean((num / den) ** 2)))


# The reference's CPU run (sample/parameter.dat values; dt_tree 2^-4, the power of two
# nearest below 2 pi / 64, because the reference's block steps need a power of two).
CFG1 = dict(R_cut0=3.0 * 2 ** (1 / 3), R_cut1=8.0, R_search0=1.1, R_search1=6.0, R_search2=1.1,
            R_search3=2.0, gamma=0.1, eta=0.01, eta_sun=0.01, eta_0=0.002, eta_sun0=0.002,
            dt_tree=2.0 ** -4, dt_min=2.0 ** -30, n_group_limit=64, theta=0.5)
