"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker, never as the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

from gplum_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_native_build():
    """Load the oracle built on this host with -O3 -march=native (oracle/_build_native;
    for the CPU baseline timing) instead of the portable build.  Must be called before
    the first use; returns False (portable build kept) if the native build fails."""
    global _LIB_PATH
    if _lib is not None:
        return _LIB_PATH.endswith(os.path.join("_build_native", "liboracle.so"))
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, timeout=300)
    except (OSError, subprocess.SubprocessError):
        return False
    _LIB_PATH = os.path.join(_HERE, "_build_native", "liboracle.so")
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        P = ctypes.POINTER(abi.Params)
        L.orc_params_default.argtypes = [P]
        L.orc_cbrt_exact_ops.argtypes = [f64]
        L.orc_cbrt_exact_ops.restype = f64
        L.orc_exp_exact_ops.argtypes = [f64]
        L.orc_exp_exact_ops.restype = f64
        L.orc_params_set_gamma.argtypes = [P, f64]
        for name in ("orc_epep_f32", "orc_epsp_quad_f32", "orc_epsp_mono_f32"):
            getattr(L, name).argtypes = [vp, i32, vp, i32, vp, f32]
        L.orc_epep_f64.argtypes = [vp, i32, vp, i32, f64, vp, vp, vp]
        L.orc_epsp_quad_f64.argtypes = [vp, i32, vp, i32, f64, vp, vp]
        L.orc_epsp_mono_f64.argtypes = [vp, i32, vp, i32, f64, vp, vp]
        for name in ("orc_cutoff_W", "orc_cutoff_K"):
            getattr(L, name).argtypes = [P, f64, f64]
            getattr(L, name).restype = f64
        L.orc_cutoff_f.argtypes = [P, f64]
        L.orc_cutoff_f.restype = f64
        L.orc_cutoff_Kx.argtypes = [f64]
        L.orc_cutoff_Kx.restype = f64
        L.orc_cutoff_dKdr.argtypes = [P, f64, f64]
        L.orc_cutoff_dKdr.restype = f64
        L.orc_cutoff_dKdt.argtypes = [P, f64, f64, f64]
        L.orc_cutoff_dKdt.restype = f64
        L.orc_correct_force_long.argtypes = [P, vp, i32, i32, vp, vp, vp]
        L.orc_neighbor_cluster.argtypes = [vp, i32, vp]
        L.orc_make_cluster_list.argtypes = [vp, i32, vp, vp]
        L.orc_make_cluster_list.restype = i32
        L.orc_star_gravity.argtypes = [P, vp]
        L.orc_calc_dt_initial.argtypes = [P, vp]
        L.orc_hermite_cluster.argtypes = [P, vp, i32, vp, vp]
        L.orc_hermite_cluster.restype = i32
        L.orc_hermite_cluster_log.argtypes = [P, vp, i32, vp, vp, vp, i32, vp]
        L.orc_hermite_cluster_log.restype = i32
        L.orc_kepler_isolated.argtypes = [P, vp, f64, f64]
        L.orc_hermite_isolated.argtypes = [P, vp]
        L.orc_cluster_prepare.argtypes = [P, vp, i32]
        L.orc_solve_kepler.argtypes = [f64, f64]
        L.orc_solve_kepler.restype = f64
        L.orc_solve_kepler_int_abs.argtypes = [f64, f64]
        L.orc_solve_kepler_int_abs.restype = f64
        L.orc_isolated.argtypes = [P, vp, f64, f64]
        L.orc_kepler_isolated_int_abs.argtypes = [P, vp, f64, f64]
        L.orc_eccentricity.argtypes = [P, vp, vp]
        L.orc_eccentricity.restype = f64
        L.orc_set_cutoff_radii.argtypes = [P, vp, i32]
        L.orc_energy.argtypes = [vp, i32, vp]
        L.orc_tree_build.argtypes = [P, vp, i32, vp]
        L.orc_tree_free.argtypes = [vp]
        L.orc_soft_force.argtypes = [P, vp, i32, vp, i32, vp]
        L.orc_soft_force.restype = ctypes.c_int64
        L.orc_nlist_free.argtypes = [vp]
        L.orc_p3t_init.argtypes = [P, vp, i32, vp, f64]
        L.orc_p3t_step.argtypes = [P, vp, vp, vp, vp, f64, vp]
        L.orc_p3t_step.restype = i32
        L.orc_loop_default.argtypes = [P, f64, vp]
        L.orc_restart_write.argtypes = [ctypes.c_char_p, ctypes.c_int32, f64, vp, vp, vp, i32]
        L.orc_restart_write.restype = i32
        L.orc_restart_read.argtypes = [ctypes.c_char_p, vp, vp, vp, vp, vp, vp, i32]
        L.orc_restart_read.restype = i32
        _lib = L
    return _lib


class NList(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("offset", ctypes.POINTER(ctypes.c_int32)),
                ("index", ctypes.POINTER(ctypes.c_int32))]

    def to_numpy(self):
        off = np.ctypeslib.as_array(self.offset, shape=(self.n + 1,)).copy()
        idx = np.ctypeslib.as_array(self.index, shape=(max(int(off[-1]), 1),))[: off[-1]].copy()
        return off, idx

    def free(self):
        lib().orc_nlist_free(ctypes.byref(self))


class Tree(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("n_node", ctypes.c_int32),
                ("origin", ctypes.c_double * 3), ("size", ctypes.c_double),
                ("key", ctypes.POINTER(ctypes.c_uint64)), ("perm", ctypes.POINTER(ctypes.c_int32))] + \
               [(n, ctypes.POINTER(ctypes.c_int32)) for n in (
                   "node_start", "node_count", "node_child", "node_nchild", "node_level")] + \
               [(n, ctypes.POINTER(ctypes.c_double)) for n in (
                   "node_mass", "node_com", "node_quad", "node_bmin", "node_bmax", "node_rs",
                   "node_center")] + \
               [("n_group", ctypes.c_int32)] + \
               [(n, ctypes.POINTER(ctypes.c_int32)) for n in (
                   "group_node", "ep_offset", "ep_leaf", "sp_offset", "sp_node")]

    def arr(self, name, n, dtype=np.int32):
        p = getattr(self, name)
        return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy()

    def lists(self):
        """(groups[G,3] = node,start,count), ep_offset, ep_leaf, sp_offset, sp_node, perm"""
        G = self.n_group
        groups = self.arr("group_node", 3 * G).reshape(G, 3)
        epo = self.arr("ep_offset", G + 1)
        spo = self.arr("sp_offset", G + 1)
        return (groups, epo, self.arr("ep_leaf", int(epo[-1])), spo,
                self.arr("sp_node", int(spo[-1])), self.arr("perm", self.n))

    def free(self):
        lib().orc_tree_free(ctypes.byref(self))


def params(**kw):
    p = abi.Params()
    lib().orc_params_default(ctypes.byref(p))
    gamma = kw.pop("gamma", None)
    for k, v in kw.items():
        setattr(p, k, v)
    if gamma is not None:
        lib().orc_params_set_gamma(ctypes.byref(p), gamma)
    return p


def epep_f32(epi, epj, force, eps2=0.0):
    lib().orc_epep_f32(abi.ptr(epi), len(epi), abi.ptr(epj), len(epj), abi.ptr(force), eps2)


def epsp_quad_f32(epi, spj, force, eps2=0.0):
    lib().orc_epsp_quad_f32(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj), abi.ptr(force), eps2)


def epsp_mono_f32(epi, spj, force, eps2=0.0):
    lib().orc_epsp_mono_f32(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj), abi.ptr(force), eps2)


def epep_f64(epi, epj, eps2=0.0, force=None):
    acc = np.zeros((len(epi), 3))
    phi = np.zeros(len(epi))
    lib().orc_epep_f64(abi.ptr(epi), len(epi), abi.ptr(epj), len(epj), eps2, abi.ptr(acc),
                       abi.ptr(phi), abi.ptr(force))
    return acc, phi


def epsp_quad_f64(epi, spj, eps2=0.0):
    acc = np.zeros((len(epi), 3))
    phi = np.zeros(len(epi))
    lib().orc_epsp_quad_f64(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj), eps2, abi.ptr(acc),
                            abi.ptr(phi))
    return acc, phi


def epsp_mono_f64(epi, spj, eps2=0.0):
    acc = np.zeros((len(epi), 3))
    phi = np.zeros(len(epi))
    lib().orc_epsp_mono_f64(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj), eps2, abi.ptr(acc),
                            abi.ptr(phi))
    return acc, phi


def set_cutoff_radii(p, pp):
    lib().orc_set_cutoff_radii(ctypes.byref(p), abi.ptr(pp), len(pp))


def energy(pp):
    out = np.zeros(5)
    lib().orc_energy(abi.ptr(pp), len(pp), abi.ptr(out))
    return out


def tree_build(p, pp):
    t = Tree()
    lib().orc_tree_build(ctypes.byref(p), abi.ptr(pp), len(pp), ctypes.byref(t))
    return t


def soft_force(p, pp, tree, use_f64=False):
    nsp = ctypes.c_int64(0)
    nep = lib().orc_soft_force(ctypes.byref(p), abi.ptr(pp), len(pp), ctypes.byref(tree),
                               int(use_f64), ctypes.byref(nsp))
    return nep, nsp.value


def correct_force_long(p, pp, initial):
    nl = NList()
    lib().orc_correct_force_long(ctypes.byref(p), abi.ptr(pp), len(pp), int(initial),
                                 ctypes.byref(nl), None, None)
    off, idx = nl.to_numpy()
    nl.free()
    return off, idx


def star_gravity(p, pp):
    for i in range(len(pp)):
        lib().orc_star_gravity(ctypes.byref(p), ctypes.c_void_p(pp.ctypes.data + i * pp.itemsize))


def hermite_cluster(p, pp):
    ed = ctypes.c_double(0.0)
    nb = ctypes.c_int64(0)
    ncol = lib().orc_hermite_cluster(ctypes.byref(p), abi.ptr(pp), len(pp), ctypes.byref(ed),
                                     ctypes.byref(nb))
    return ncol, ed.value, nb.value


def hermite_cluster_log(p, pp, cap=256):
    """hermite_cluster with the collision log's records: (n_col, edisp, block_steps,
    records[n_col, 20]) (fields COLREC_FIELDS)."""
    rec = np.zeros((cap, 20), dtype=np.float64)
    ed = ctypes.c_double(0.0)
    nb = ctypes.c_int64(0)
    nrec = ctypes.c_int(0)
    ncol = lib().orc_hermite_cluster_log(ctypes.byref(p), abi.ptr(pp), len(pp), ctypes.byref(ed),
                                         ctypes.byref(nb), ctypes.c_void_p(rec.ctypes.data), cap,
                                         ctypes.byref(nrec))
    return ncol, ed.value, nb.value, rec[: min(nrec.value, cap)].copy()


def cluster_prepare(p, pp, offsets):
    """Sun + hard acc/jerk and initial block steps for every cluster (in place)."""
    for c in range(len(offsets) - 1):
        a, b = int(offsets[c]), int(offsets[c + 1])
        lib().orc_cluster_prepare(ctypes.byref(p), ctypes.c_void_p(pp.ctypes.data + a * pp.itemsize), b - a)


def hermite_clusters(p, pp, offsets):
    """orc_hermite_cluster over every cluster (in place); returns (n_col, edisp, block_steps)."""
    nc, ed, nb = 0, 0.0, 0
    for c in range(len(offsets) - 1):
        a, b = int(offsets[c]), int(offsets[c + 1])
        sub = pp[a:b].copy()
        r = hermite_cluster(p, sub)
        pp[a:b] = sub
        nc += r[0]
        ed += r[1]
        nb += r[2]
    return nc, ed, nb


def kepler_isolated(p, pp, t0, t1, int_abs=False):
    """timeIntegrateKepler_isolated on every particle; int_abs: solveKeplerEq with the
    int abs(int) outcome of kepler.h:23 (Newton only for residuals of 1 or more)."""
    f = lib().orc_kepler_isolated_int_abs if int_abs else lib().orc_kepler_isolated
    for i in range(len(pp)):
        f(ctypes.byref(p), ctypes.c_void_p(pp.ctypes.data + i * pp.itemsize), t0, t1)


def hermite_isolated(p, pp):
    """timeIntegrate_isolated (Sun-only Hermite over one soft step) on every particle."""
    for i in range(len(pp)):
        lib().orc_hermite_isolated(ctypes.byref(p), ctypes.c_void_p(pp.ctypes.data + i * pp.itemsize))


def isolated(p, pp, t0, t1):
    """The isolated-particle branch of timeIntegrate (Kepler or Sun-only Hermite)."""
    for i in range(len(pp)):
        lib().orc_isolated(ctypes.byref(p), ctypes.c_void_p(pp.ctypes.data + i * pp.itemsize), t0, t1)


def eccentricity(p, pp):
    return np.array([lib().orc_eccentricity(ctypes.byref(p), ctypes.c_void_p(pp.ctypes.data + i * pp.itemsize),
                                            None) for i in range(len(pp))])


class Loop(ctypes.Structure):
    """orc_loop: main-loop state (boundary, reset interval, step counter, outputs)."""
    _fields_ = [("r_max", ctypes.c_double), ("r_min", ctypes.c_double), ("reset_step", ctypes.c_int32),
                ("istep", ctypes.c_int32), ("n_remove", ctypes.c_int32), ("recomputed", ctypes.c_int32),
                ("e_now", ctypes.c_double * 5), ("removed", ctypes.c_void_p), ("removed_cap", ctypes.c_int32),
                ("n_col_rec", ctypes.c_int32), ("col_rec", ctypes.c_void_p), ("col_cap", ctypes.c_int32)]


# the collision log's record (orc_hermite_cluster_log, collision_record in orc_hard.c)
COLREC_FIELDS = ["time", "id_imp", "id_tar", "mass_imp", "mass_tar"] + \
    ["%s_%s_%s" % (q, w, c) for q, w in (("pos", "imp"), ("pos", "tar"), ("vel", "imp"), ("vel", "tar"))
     for c in "xyz"] + ["dpos_g", "dvel_g", "pad"]


class P3T:
    """Whole-system oracle run (initialisation + soft steps)."""

    def __init__(self, p, pp, time_sys=0.0, r_max=None, r_min=None, reset_step=None):
        self.p = p
        self.pp = pp.copy()
        self.n = ctypes.c_int(len(pp))
        self.nl = NList()
        self.edisp = ctypes.c_double(0.0)
        self.time = time_sys
        self.loop = Loop()
        lib().orc_loop_default(ctypes.byref(p), time_sys, ctypes.byref(self.loop))
        for k, v in (("r_max", r_max), ("r_min", r_min), ("reset_step", reset_step)):
            if v is not None:
                setattr(self.loop, k, v)
        self.removed_buf = np.zeros(4096, dtype=abi.PARTICLE)
        self.loop.removed = self.removed_buf.ctypes.data
        self.loop.removed_cap = len(self.removed_buf)
        self.removed = []          # per step: the removed particles (ascending id)
        self.col_buf = np.zeros((4096, 20), dtype=np.float64)
        self.loop.col_rec = self.col_buf.ctypes.data
        self.loop.col_cap = len(self.col_buf)
        self.collisions = []       # per step: the collision records (cluster, then time order)
        lib().orc_p3t_init(ctypes.byref(p), abi.ptr(self.pp), len(self.pp), ctypes.byref(self.nl),
                           time_sys)

    def step(self):
        ncol = lib().orc_p3t_step(ctypes.byref(self.p), abi.ptr(self.pp), ctypes.byref(self.n),
                                  ctypes.byref(self.nl), ctypes.byref(self.edisp), self.time,
                                  ctypes.byref(self.loop))
        self.time += self.p.dt_tree
        self.removed.append(self.removed_buf[: min(self.loop.n_remove, len(self.removed_buf))].copy())
        self.collisions.append(self.col_buf[: self.loop.n_col_rec].copy())
        if self.n.value != len(self.pp):
            self.pp = self.pp[: self.n.value].copy()
        return ncol

    def neighbor_graph(self):
        return self.nl.to_numpy()

    def close(self):
        self.nl.free()
