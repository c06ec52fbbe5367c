"""Numpy/ctypes mirrors of the structs in include/gplum_gpu.h and include/gplum_system.h.

The layouts are the reference's individual-cutoff build (src/particle.h:16-170) and the
FDPS super-particle records; `check_layouts()` asserts the byte sizes the C side
static_asserts.
"""
import ctypes

import numpy as np

NEIGHBOR_INFO = np.dtype([("number", "<i4"), ("rank", "<i4"), ("id_max", "<i4"), ("id_min", "<i4")])

FORCE = np.dtype([("acc", "<f4", 3), ("phi", "<f4"), ("neighbor", NEIGHBOR_INFO)], align=True)

EPI = np.dtype([("id_local", "<i4"), ("myrank", "<i4"), ("pos", "<f8", 3),
                ("r_out", "<f8"), ("r_search", "<f8")], align=True)

EPJ = np.dtype([("id_local", "<i4"), ("myrank", "<i4"), ("pos", "<f8", 3),
                ("r_out", "<f8"), ("r_search", "<f8"), ("id", "<i8"), ("mass", "<f8"),
                ("vel", "<f8", 3), ("acc_d", "<f8", 3)], align=True)

SPJ_MONO = np.dtype([("mass", "<f8"), ("pos", "<f8", 3)], align=True)

SPJ_QUAD = np.dtype([("mass", "<f8"), ("pos", "<f8", 3), ("quad", "<f8", 6)], align=True)

PARTICLE = np.dtype([
    ("id", "<i8"), ("id_cluster", "<i8"),
    ("mass", "<f8"), ("r_planet", "<f8"), ("f", "<f8"),
    ("pos", "<f8", 3), ("vel", "<f8", 3), ("acc", "<f8", 3),
    ("acc_d", "<f8", 3), ("jerk_d", "<f8", 3), ("acc_s", "<f8", 3), ("jerk_s", "<f8", 3),
    ("phi", "<f8"), ("phi_d", "<f8"), ("phi_s", "<f8"),
    ("r_out", "<f8"), ("r_out_inv", "<f8"), ("r_search", "<f8"), ("v_disp", "<f8"),
    ("time", "<f8"), ("dt", "<f8"), ("acc0", "<f8"),
    ("neighbor", NEIGHBOR_INFO), ("n_cluster", "<i4"), ("flags", "<i4"),
], align=True)

FLAG_DEAD = 1
FLAG_MERGED = 2


class Params(ctypes.Structure):
    """gplum_params (FPGrav static parameters + tree parameters)."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "m_sun", "dens", "eps2", "eps2_sun",
        "R_cut0", "R_cut1", "R_search0", "R_search1", "R_search2", "R_search3",
        "gamma", "g_1_inv", "g_1_inv7", "w_y", "f1",
        "r_cut_min", "r_cut_max", "p_cut", "increase_factor",
        "dt_tree", "dt_min", "eta", "eta_0", "eta_sun", "eta_sun0", "alpha2",
        "theta")] + [(n, ctypes.c_int32) for n in (
            "n_leaf_limit", "n_group_limit", "quadrupole", "collision")]


def check_layouts():
    assert NEIGHBOR_INFO.itemsize == 16
    assert FORCE.itemsize == 32
    assert EPI.itemsize == 48
    assert EPJ.itemsize == 112
    assert SPJ_MONO.itemsize == 32
    assert SPJ_QUAD.itemsize == 80
    assert ctypes.sizeof(Params) == 27 * 8 + 4 * 4


def cleared_force(n):
    """ForceGrav::clear() values (src/particle.h:48-53, 81-85)."""
    f = np.zeros(n, dtype=FORCE)
    f["neighbor"]["id_max"] = -1
    f["neighbor"]["id_min"] = 2 ** 31 - 1  # S32_MAX-1, S32_MAX = 2^31 (main_p3t.cpp:32)
    return f


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)
