"""Synthetic initial conditions for the planetesimal-disk workloads.

The disk law restates SolidDisk::createInitialCondition (src/disk.h:77-171): semi-major
axes with dN/da ~ a^(1-p) between a_in and a_out, Keplerian orbits with random angles,
eccentricity/inclination scaled by the reduced Hill radius h = (m/3 M_sun)^(1/3).  The
reference draws |N(0, sigma)| (src/mathfunc.h:25-49); the workloads here ask for Rayleigh
distributions with a given rms, so sigma_R = rms/sqrt(2).  Random numbers come from numpy's
PCG64 (seeded), not drand48, so samples follow the same law but are not the reference's.

Units: G = M_sun = AU = 1, time unit yr/2pi (src/main_p3t.cpp:25-31).
"""
import math

import numpy as np

from . import abi

L_CGS = 14959787070000.0
M_CGS = 1.9884e33


def density_code(rho_cgs):
    """g/cm^3 -> M_sun/AU^3 (getvalue(.., M_CGS/L_CGS^3), src/particle.h:513)."""
    return rho_cgs / (M_CGS / (L_CGS ** 3))


def solve_kepler(l, e, iters=30):
    """Newton iteration for u - e sin u = l (vectorised); stops once the largest
    update is at the rounding level of u (|u| <= 2 pi)."""
    u = l + e * np.sin(l)
    for _ in range(iters):
        f = u - e * np.sin(u) - l
        un = u - f / (1.0 - e * np.cos(u))
        done = np.max(np.abs(un - u)) <= 4e-15
        u = un
        if done:
            break
    return u


def elements_to_posvel(ax, ecc, inc, l, omg, OMG, m_sun=1.0):
    """orbitalElement2PosVel with P, Q from the angles (src/disk.h:150-166, kepler.h:77-93)."""
    u = solve_kepler(l, ecc)
    n = np.sqrt(m_sun / ax ** 3)
    P = np.stack([np.cos(omg) * np.cos(OMG) - np.sin(omg) * np.sin(OMG) * np.cos(inc),
                  np.cos(omg) * np.sin(OMG) + np.sin(omg) * np.cos(OMG) * np.cos(inc),
                  np.sin(omg) * np.sin(inc)], axis=1)
    Q = np.stack([-np.sin(omg) * np.cos(OMG) - np.cos(omg) * np.sin(OMG) * np.cos(inc),
                  -np.sin(omg) * np.sin(OMG) + np.cos(omg) * np.cos(OMG) * np.cos(inc),
                  np.cos(omg) * np.sin(inc)], axis=1)
    cosu, sinu = np.cos(u)[:, None], np.sin(u)[:, None]
    esq = np.sqrt(1.0 - ecc * ecc)[:, None]
    pos = ax[:, None] * ((cosu - ecc[:, None]) * P + esq * sinu * Q)
    rinv = 1.0 / np.sqrt((pos * pos).sum(1))[:, None]
    vel = (ax * ax * n)[:, None] * rinv * (-sinu * P + esq * cosu * Q)
    return pos, vel


def disk(n, mass, a_in, a_out, p=1.5, ecc_hill=2.0, inc_hill=1.0, rho_cgs=2.0, seed=0,
         m_sun=1.0, f=1.0):
    """n equal-mass planetesimals; rms e = ecc_hill*h, rms i = inc_hill*h (Rayleigh)."""
    rng = np.random.default_rng(seed)
    R = rng.random(n)
    if p != 2:
        q = 2.0 - p
        ax = ((a_out ** q - a_in ** q) * R + a_in ** q) ** (1.0 / q)
    else:
        ax = np.exp((math.log(a_out) - math.log(a_in)) * R + math.log(a_in))
    h = (mass / (3.0 * m_sun)) ** (1.0 / 3.0)
    ecc = rng.rayleigh(ecc_hill * h / math.sqrt(2.0), n)
    inc = rng.rayleigh(inc_hill * h / math.sqrt(2.0), n)
    l = 2 * math.pi * rng.random(n)
    omg = 2 * math.pi * rng.random(n)
    OMG = 2 * math.pi * rng.random(n)
    pos, vel = elements_to_posvel(ax, ecc, inc, l, omg, OMG, m_sun)
    pp = np.zeros(n, dtype=abi.PARTICLE)
    pp["id"] = np.arange(n)
    pp["id_cluster"] = np.arange(n)
    pp["mass"] = mass
    pp["pos"] = pos
    pp["vel"] = vel
    pp["f"] = f
    pp["r_planet"] = (0.75 * mass / (math.pi * density_code(rho_cgs))) ** (1.0 / 3.0)
    return pp


def add_embryos(pp, n_emb, mass, a_in, a_out, rho_cgs=2.0, seed=7):
    """Embryos on circular, coplanar orbits evenly spaced in a (random phases)."""
    rng = np.random.default_rng(seed)
    ax = np.linspace(a_in, a_out, n_emb + 2)[1:-1]
    z = np.zeros(n_emb)
    pos, vel = elements_to_posvel(ax, z, z, 2 * math.pi * rng.random(n_emb), z, z)
    e = np.zeros(n_emb, dtype=abi.PARTICLE)
    e["id"] = pp["id"].max() + 1 + np.arange(n_emb)
    e["id_cluster"] = e["id"]
    e["mass"] = mass
    e["pos"] = pos
    e["vel"] = vel
    e["f"] = 1.0
    e["r_planet"] = (0.75 * mass / (math.pi * density_code(rho_cgs))) ** (1.0 / 3.0)
    return np.concatenate([pp, e])


def clusters(n_clusters, seed=2, size_min=2, size_max=256, m_lo=1e-12, m_hi=1e-8, m_ref=1e-10,
             sigma_hill=2.0, r_cut0=3.0 * 2.0 ** (1.0 / 3.0), rho_cgs=2.0):
    """Independent compact groups around a circular 1 AU orbit (batched-hard workload).

    Sizes n = floor(exp(U[ln size_min, ln(size_max+1)))); member offsets N(0, sigma) in 3-D
    with sigma = sigma_hill * r_H(m_ref); velocities circular Keplerian at each member's
    position; masses log-uniform in [m_lo, m_hi].  r_out = r_cut0 * r_H(m_i) at 1 AU.
    Returns (particles, offsets[n_clusters+1]).
    """
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(math.log(size_min), math.log(size_max + 1), n_clusters)))
    sizes = np.clip(sizes.astype(np.int64), size_min, size_max)
    off = np.zeros(n_clusters + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    phase = rng.uniform(0, 2 * math.pi, n_clusters)
    centre = np.stack([np.cos(phase), np.sin(phase), np.zeros(n_clusters)], 1)
    sigma = sigma_hill * (m_ref / 3.0) ** (1.0 / 3.0)
    cid = np.repeat(np.arange(n_clusters), sizes)
    pos = centre[cid] + rng.normal(0.0, sigma, (n, 3))
    rxy = np.sqrt(pos[:, 0] ** 2 + pos[:, 1] ** 2)
    vk = 1.0 / np.sqrt(rxy)
    vel = np.stack([-pos[:, 1] / rxy * vk, pos[:, 0] / rxy * vk, np.zeros(n)], 1)
    mass = np.exp(rng.uniform(math.log(m_lo), math.log(m_hi), n))
    pp = np.zeros(n, dtype=abi.PARTICLE)
    pp["id"] = np.arange(n)
    pp["id_cluster"] = off[cid]
    pp["mass"] = mass
    pp["pos"] = pos
    pp["vel"] = vel
    pp["f"] = 1.0
    pp["r_planet"] = (0.75 * mass / (math.pi * density_code(rho_cgs))) ** (1.0 / 3.0)
    r_out = r_cut0 * (mass / 3.0) ** (1.0 / 3.0)
    pp["r_out"] = r_out
    pp["r_out_inv"] = 1.0 / r_out
    pp["r_search"] = 1.1 * r_out
    pp["neighbor"]["number"] = np.repeat(sizes - 1, sizes)
    return pp, off


def cluster_acc0(pp, off):
    """acc0 as correctForceLong leaves it for a cluster whose members are all
    neighbours (gravity_soft.h:330-340): (n - 1) / sum_j max(r_out_i, r_out_j)^2 / m_j."""
    for c in range(len(off) - 1):
        a, b = int(off[c]), int(off[c + 1])
        ro = pp["r_out"][a:b]
        w = np.maximum.outer(ro, ro) ** 2 / pp["mass"][a:b][None, :]
        np.fill_diagonal(w, 0.0)
        pp["acc0"][a:b] = (b - a - 1) / w.sum(1)


def read_snapshot(path):
    """ASCII snapshot (FPGrav::readAscii, src/particle.h:829-843) -> (header, particles).

    Header line: time, n_tot, n_tot, etot0, ekin0, ephi_sun0, ephi_planet0, edisp0, then the
    same five energies for the current time (src/energy.h:70-141)."""
    with open(path) as fh:
        header = [float(x) for x in fh.readline().split()]
        data = np.loadtxt(fh, ndmin=2)
    n = data.shape[0]
    pp = np.zeros(n, dtype=abi.PARTICLE)
    pp["id"] = data[:, 0].astype(np.int64)
    pp["id_cluster"] = pp["id"]
    pp["mass"] = data[:, 1]
    pp["r_planet"] = data[:, 2]
    pp["f"] = data[:, 3]
    pp["pos"] = data[:, 4:7]
    pp["vel"] = data[:, 7:10]
    pp["neighbor"]["number"] = data[:, 10].astype(np.int32)
    return header, pp
