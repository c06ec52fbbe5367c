"""ctypes binding of libgplum_gpu.so (the HIP implementation).

Loading fails loudly when the library is missing: there is no CPU fallback.  Build it with
`python -m gplum_amd.build` (or __graft_entry__.build()).
"""
import ctypes
import os
import re

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgplum_gpu.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")
_lib = None


class GplumError(RuntimeError):
    pass


def declared_symbols():
    """Function names declared in include/*.h (the C ABI surface)."""
    names = []
    for h in sorted(os.listdir(INCLUDE_DIR)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE_DIR, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(gplum_[a-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GplumError("libgplum_gpu.so is not built (%s); run python -m gplum_amd.build"
                             % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        P = ctypes.POINTER(abi.Params)
        L.gplum_last_error.restype = ctypes.c_char_p
        L.gplum_params_default.argtypes = [P]
        L.gplum_params_set_gamma.argtypes = [P, f64]
        L.gplum_device_info.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(i32),
                                        ctypes.POINTER(f64)]
        for n in ("gplum_calc_force_epep", "gplum_calc_force_epsp_quad", "gplum_calc_force_epsp_mono"):
            getattr(L, n).argtypes = [vp, i32, vp, i32, vp, f32]
        L.gplum_calc_force_epep_noalloc.argtypes = [vp, i32, vp, i32, vp, f32, vp]
        for n in ("gplum_calc_force_epep_f64", "gplum_calc_force_epsp_quad_f64"):
            getattr(L, n).argtypes = [vp, i32, vp, i32, vp, f32, vp, vp]
        L.gplum_system_create.argtypes = [P, ctypes.POINTER(vp)]
        L.gplum_system_destroy.argtypes = [vp]
        L.gplum_system_set_params.argtypes = [vp, P]
        L.gplum_system_set_particles.argtypes = [vp, vp, i32, f64]
        L.gplum_system_get_particles.argtypes = [vp, vp, i32]
        L.gplum_system_count.argtypes = [vp, ctypes.POINTER(i32)]
        L.gplum_system_initialize.argtypes = [vp]
        L.gplum_system_set_radii.argtypes = [vp]
        L.gplum_system_soft_force.argtypes = [vp, i32, i32]
        L.gplum_system_hard_part.argtypes = [vp, i32]
        L.gplum_system_step.argtypes = [vp, i32]
        L.gplum_system_set_boundary.argtypes = [vp, f64, f64, i32]
        L.gplum_system_set_event_logs.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
        L.gplum_system_energy.argtypes = [vp, vp]
        L.gplum_system_neighbors.argtypes = [vp, vp, vp, i32, ctypes.POINTER(i32)]
        L.gplum_system_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.gplum_system_reset_stats.argtypes = [vp, i32]
        L.gplum_system_set_forecast.argtypes = [vp, i32, i32, i32, f32]
        L.gplum_system_set_forecast_every.argtypes = [vp, i32]
        L.gplum_system_set_variant.argtypes = [vp, i32, i32]
        L.gplum_system_get_forecast.argtypes = [vp, vp, vp, vp, ctypes.POINTER(i32)]
        L.gplum_system_put_forecast.argtypes = [vp, vp, vp, vp]
        L.gplum_hard_integrate.argtypes = [P, vp, i32, vp, i32, i32, ctypes.POINTER(f64),
                                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(f64),
                                           ctypes.POINTER(ctypes.c_int64)]
        L.gplum_mlp_forward.argtypes = [vp, vp, i32, vp]
        L.gplum_mlp_predict.argtypes = [vp, vp, i32, vp]
        L.gplum_mlp_loss_grad.argtypes = [vp, vp, vp, i32, ctypes.POINTER(f64), vp]
        L.gplum_mlp_train.argtypes = [vp, vp, vp, i32, i32, f32, ctypes.POINTER(f64),
                                      ctypes.POINTER(f64)]
        _lib = L
    return _lib


def hard_integrate(p, pp, offsets, persistent=True, prepare=False, counts=False, giant=False, lane_per_particle=False):
    """Integrate clusters pp[offsets[c]:offsets[c+1]] over one soft step (in place).
    prepare: compute acc/jerk and the first block step on the device first.
    giant: every cluster above 64 particles on the whole-GPU path (same results).
    lane_per_particle: clusters of 2 to 4 bodies on the one-lane-per-particle path instead
    of the lane-per-term path (same results).
    Returns (edisp, block_steps, kernel_ms), plus pair evaluations when counts."""
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    ed = ctypes.c_double(0.0)
    nb = ctypes.c_int64(0)
    ms = ctypes.c_double(0.0)
    npair = ctypes.c_int64(0)
    mode = (1 if persistent else 0) | (2 if prepare else 0) | (4 if giant else 0) | (8 if lane_per_particle else 0)
    check(lib().gplum_hard_integrate(ctypes.byref(p), abi.ptr(pp), len(pp), abi.ptr(off), len(off) - 1, mode,
                                     ctypes.byref(ed), ctypes.byref(nb), ctypes.byref(ms),
                                     ctypes.byref(npair)))
    if counts:
        return ed.value, nb.value, ms.value, npair.value
    return ed.value, nb.value, ms.value


def hard_integrate_graph(p, pp, offsets, replays=4, prepare=False):
    """gplum_hard_integrate_graph: the persistent launch captured into a hipGraph and
    replayed `replays` times from the same uploaded state (in place, first replay's
    result).  Returns (edisp, block_steps, ms_per_replay, mismatches)."""
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    ed = ctypes.c_double(0.0)
    nb = ctypes.c_int64(0)
    ms = ctypes.c_double(0.0)
    bad = ctypes.c_int(-1)
    check(lib().gplum_hard_integrate_graph(ctypes.byref(p), abi.ptr(pp), len(pp), abi.ptr(off), len(off) - 1,
                                           int(prepare), int(replays), ctypes.byref(ed), ctypes.byref(nb),
                                           ctypes.byref(ms), ctypes.byref(bad)))
    return ed.value, nb.value, ms.value, bad.value


def hermite_dt_selfcheck(p, n, seed=1):
    """gplum_hermite_dt_selfcheck: (mismatches, fallbacks) of the fast time-step criterion
    against the correctly rounded one on n random states."""
    m = ctypes.c_int64(-1)
    f = ctypes.c_int64(-1)
    check(lib().gplum_hermite_dt_selfcheck(ctypes.byref(p), ctypes.c_int64(n), ctypes.c_uint64(seed),
                                           ctypes.byref(m), ctypes.byref(f)))
    return m.value, f.value


class SnapshotHeader(ctypes.Structure):
    """gplum_snapshot_header (include/gplum_system.h)."""
    _fields_ = [("time", ctypes.c_double), ("n_body", ctypes.c_int32), ("id_next", ctypes.c_int32),
                ("e_init", ctypes.c_double * 5), ("e_now", ctypes.c_double * 5)]


def snapshot_read(path):
    """(header, particles) of an ASCII snapshot in the reference's format."""
    h = SnapshotHeader()
    n = ctypes.c_int(0)
    check(lib().gplum_snapshot_read(path.encode(), ctypes.byref(h), None, 0, ctypes.byref(n)))
    pp = np.zeros(n.value, dtype=abi.PARTICLE)
    check(lib().gplum_snapshot_read(path.encode(), ctypes.byref(h), abi.ptr(pp), len(pp), ctypes.byref(n)))
    return h, pp


def snapshot_write(path, header, pp):
    pp = np.ascontiguousarray(pp)
    check(lib().gplum_snapshot_write(path.encode(), ctypes.byref(header), abi.ptr(pp), len(pp)))


ENERGY_FIELDS = ("etot", "ekin", "ephi_sun", "ephi_planet", "ephi", "ephi_d", "edisp")


class Energy(ctypes.Structure):
    """gplum_energy (Energy, src/energy.h:3-11)."""
    _fields_ = [(n, ctypes.c_double) for n in ENERGY_FIELDS]


class FileHeader(ctypes.Structure):
    """gplum_file_header (FileHeader, src/energy.h:70-79), the binary restart header."""
    _fields_ = [("n_body", ctypes.c_int32), ("id_next", ctypes.c_int32), ("time", ctypes.c_double),
                ("e_init", Energy), ("e_now", Energy)]


def restart_write(path, header, pp):
    """The binary restart file snap_tmp.dat (makeSnapTmp) of particles pp."""
    pp = np.ascontiguousarray(pp, dtype=abi.PARTICLE)
    check(lib().gplum_restart_write(path.encode(), ctypes.byref(header), abi.ptr(pp), len(pp)))


def restart_read(path):
    """(header, particles) of a binary restart file."""
    h = FileHeader()
    n = ctypes.c_int(0)
    check(lib().gplum_restart_read(path.encode(), ctypes.byref(h), None, 0, ctypes.byref(n)))
    pp = np.zeros(n.value, dtype=abi.PARTICLE)
    check(lib().gplum_restart_read(path.encode(), ctypes.byref(h), abi.ptr(pp), len(pp), ctypes.byref(n)))
    return h, pp


class Stats(ctypes.Structure):
    """gplum_stats (include/gplum_system.h)."""
    _fields_ = [(n, ctypes.c_int64) for n in (
        "steps", "n_epep_near", "n_epep_far", "n_epsp", "n_sp_entries", "n_groups", "n_candidates",
        "n_neighbors",
        "n_clusters", "n_cluster_particles", "max_cluster", "block_steps", "n_collisions",
        "cand_retries", "n_force_calls", "fc_calls", "fc_slots", "fc_ovf_groups",
        "fc_ovf_candidates", "hard_pairs", "max_block_steps", "max_block_cluster", "n_force_f64_calls",
        "n_removed", "n_recomputes")] + [(n, ctypes.c_double) for n in (
            "force_ms", "tree_ms", "finalize_ms", "hard_ms", "time_sys", "fc_loss", "fc_ms",
            "hard_kernel_ms")] + [("n_hard_launches", ctypes.c_int64), ("fc_train_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    """gplum_tuning (include/gplum_system.h)."""
    _fields_ = [("n_group_limit", ctypes.c_int32), ("soft_tile", ctypes.c_int32),
                ("hard_wg_per_cu", ctypes.c_int32), ("hard_lds_max", ctypes.c_int32),
                ("hard_giant_min", ctypes.c_int32), ("tree_coop_min", ctypes.c_int32),
                ("hard_terms_max", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("soft_ms", ctypes.c_double), ("hard_ms", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class System:
    """Device-resident P3T system (include/gplum_system.h)."""

    def __init__(self, p, particles=None, time_sys=0.0):
        self.p = p
        self.h = ctypes.c_void_p()
        check(lib().gplum_system_create(ctypes.byref(p), ctypes.byref(self.h)))
        if particles is not None:
            self.set_particles(particles, time_sys)

    def close(self):
        if self.h:
            lib().gplum_system_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        self.p = p
        check(lib().gplum_system_set_params(self.h, ctypes.byref(p)))

    def set_particles(self, pp, time_sys=0.0):
        pp = np.ascontiguousarray(pp, dtype=abi.PARTICLE)
        check(lib().gplum_system_set_particles(self.h, abi.ptr(pp), len(pp), time_sys))

    def count(self):
        n = ctypes.c_int(0)
        check(lib().gplum_system_count(self.h, ctypes.byref(n)))
        return n.value

    def particles(self):
        n = self.count()
        pp = np.zeros(n, dtype=abi.PARTICLE)
        check(lib().gplum_system_get_particles(self.h, abi.ptr(pp), n))
        return pp

    def initialize(self):
        check(lib().gplum_system_initialize(self.h))

    def set_radii(self):
        check(lib().gplum_system_set_radii(self.h))

    def soft_force(self, initial=False, use_f64=False):
        check(lib().gplum_system_soft_force(self.h, int(initial), int(use_f64)))

    def hard_part(self, persistent=True):
        check(lib().gplum_system_hard_part(self.h, int(persistent)))

    def step(self, nsteps=1):
        check(lib().gplum_system_step(self.h, nsteps))

    def set_boundary(self, r_max=40.0, r_min=0.1, reset_step=1024):
        check(lib().gplum_system_set_boundary(self.h, r_max, r_min, reset_step))

    def set_event_logs(self, collision_path=None, remove_path=None):
        enc = lambda x: x.encode() if x else None
        check(lib().gplum_system_set_event_logs(self.h, enc(collision_path), enc(remove_path)))

    def energy(self):
        out = np.zeros(6)
        check(lib().gplum_system_energy(self.h, abi.ptr(out)))
        return out

    def neighbors(self):
        """(offset[n+1], index) of the neighbour graph in the current device order."""
        n = self.count()
        off = np.zeros(n + 1, dtype=np.int32)
        tot = ctypes.c_int(0)
        check(lib().gplum_system_neighbors(self.h, abi.ptr(off), None, 0, ctypes.byref(tot)))
        idx = np.zeros(max(tot.value, 1), dtype=np.int32)
        check(lib().gplum_system_neighbors(self.h, abi.ptr(off), abi.ptr(idx), len(idx),
                                           ctypes.byref(tot)))
        return off, idx[: tot.value]

    def stats(self):
        s = Stats()
        check(lib().gplum_system_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self, timing=True):
        check(lib().gplum_system_reset_stats(self.h, int(timing)))

    def set_variant(self, fp64_soft=False, per_cluster_hard=False):
        check(lib().gplum_system_set_variant(self.h, int(fp64_soft), int(per_cluster_hard)))

    def restart_write(self, path, id_next=0):
        """makeSnapTmp: the system's particles and energies as a binary restart file."""
        check(lib().gplum_system_restart_write(self.h, path.encode(), int(id_next)))

    def snapshot(self, path, id_next=0):
        """makeSnap: the system's particles and energies as a reference-format snapshot."""
        check(lib().gplum_system_snapshot(self.h, path.encode(), int(id_next)))

    def energy_line(self):
        buf = ctypes.create_string_buffer(256)
        check(lib().gplum_system_energy_line(self.h, buf, len(buf)))
        return buf.value.decode()

    def autotune(self, trials=2):
        """Pick the launch shapes (i-group size, soft j-tile, Hermite grid and LDS
        binning threshold) by timing them on the current state; returns the choice."""
        t = Tuning()
        check(lib().gplum_system_autotune(self.h, int(trials), ctypes.byref(t)))
        return t.as_dict()

    def get_tuning(self):
        t = Tuning()
        check(lib().gplum_system_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, n_group_limit=256, soft_tile=256, hard_wg_per_cu=8, hard_lds_max=256,
                   hard_giant_min=1024, tree_coop_min=0, hard_terms_max=4, **_):
        t = Tuning(n_group_limit, soft_tile, hard_wg_per_cu, hard_lds_max, hard_giant_min, tree_coop_min,
                   hard_terms_max, 0, 0.0, 0.0)
        check(lib().gplum_system_set_tuning(self.h, ctypes.byref(t)))

    FORECAST_MODES = {"off": 0, "network": 1, "uniform": 2}

    def set_forecast(self, mode="network", train_steps=2, batch=8192, lr=3e-3):
        check(lib().gplum_system_set_forecast(self.h, self.FORECAST_MODES[mode], train_steps, batch, lr))

    def set_forecast_every(self, every):
        """A fresh forecast and training round every `every` soft forces."""
        check(lib().gplum_system_set_forecast_every(self.h, int(every)))

    def get_forecast(self):
        """(theta, mu, inv_sd, trained_rounds) of the system's forecaster network."""
        theta = np.zeros(4545, np.float32)
        mu = np.zeros(4, np.float32)
        isd = np.zeros(4, np.float32)
        tr = ctypes.c_int(0)
        check(lib().gplum_system_get_forecast(self.h, abi.ptr(theta), abi.ptr(mu), abi.ptr(isd),
                                              ctypes.byref(tr)))
        return theta, mu, isd, tr.value

    def put_forecast(self, theta, mu, inv_sd):
        theta, mu, inv_sd = (np.ascontiguousarray(a, np.float32) for a in (theta, mu, inv_sd))
        check(lib().gplum_system_put_forecast(self.h, abi.ptr(theta), abi.ptr(mu), abi.ptr(inv_sd)))


def neighbor_sets_by_id(pp, off, idx):
    """{id_i: sorted tuple of neighbour ids} from a CSR graph over array pp."""
    ids = pp["id"]
    return {int(ids[i]): tuple(sorted(int(ids[j]) for j in idx[off[i]:off[i + 1]]))
            for i in range(len(pp))}


def check(rc):
    if rc != 0:
        raise GplumError("gplum error %d: %s" % (rc, lib().gplum_last_error().decode()))


def params(**kw):
    p = abi.Params()
    lib().gplum_params_default(ctypes.byref(p))
    gamma = kw.pop("gamma", None)
    for k, v in kw.items():
        setattr(p, k, v)
    if gamma is not None:
        lib().gplum_params_set_gamma(ctypes.byref(p), gamma)
    return p


def device_info():
    name = ctypes.create_string_buffer(64)
    cu = ctypes.c_int(0)
    clk = ctypes.c_double(0.0)
    check(lib().gplum_device_info(name, 64, ctypes.byref(cu), ctypes.byref(clk)))
    return name.value.decode(), cu.value, clk.value


def calc_force_epep(epi, epj, force, eps2=0.0):
    check(lib().gplum_calc_force_epep(abi.ptr(epi), len(epi), abi.ptr(epj), len(epj),
                                      abi.ptr(force), eps2))


def calc_force_epsp_quad(epi, spj, force, eps2=0.0):
    check(lib().gplum_calc_force_epsp_quad(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj),
                                           abi.ptr(force), eps2))


def calc_force_epsp_mono(epi, spj, force, eps2=0.0):
    check(lib().gplum_calc_force_epsp_mono(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj),
                                           abi.ptr(force), eps2))


def calc_force_epep_f64(epi, epj, force, eps2=0.0):
    acc = np.zeros((len(epi), 3))
    phi = np.zeros(len(epi))
    check(lib().gplum_calc_force_epep_f64(abi.ptr(epi), len(epi), abi.ptr(epj), len(epj),
                                          abi.ptr(force), eps2, abi.ptr(acc), abi.ptr(phi)))
    return acc, phi


def calc_force_epsp_quad_f64(epi, spj, force, eps2=0.0):
    acc = np.zeros((len(epi), 3))
    phi = np.zeros(len(epi))
    check(lib().gplum_calc_force_epsp_quad_f64(abi.ptr(epi), len(epi), abi.ptr(spj), len(spj),
                                               abi.ptr(force), eps2, abi.ptr(acc), abi.ptr(phi)))
    return acc, phi


# ------------------------------------------------------------ forecaster network

def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mlp_forward(theta, x):
    """rate[i] for features x[n, 4] (converted to the kernels' feature-major layout)."""
    theta, xt = _f32(theta), _f32(np.asarray(x).T)
    n = xt.shape[1]
    out = np.zeros(n, np.float32)
    check(lib().gplum_mlp_forward(abi.ptr(theta), abi.ptr(xt), n, abi.ptr(out)))
    return out


def mlp_predict(theta, x):
    """The sizing forecast: mlp_forward's rates with the hidden layer on MFMA (bf16)."""
    theta, xt = _f32(theta), _f32(np.asarray(x).T)
    n = xt.shape[1]
    out = np.zeros(n, np.float32)
    check(lib().gplum_mlp_predict(abi.ptr(theta), abi.ptr(xt), n, abi.ptr(out)))
    return out


def mlp_loss_grad(theta, x, y):
    theta, xt, y = _f32(theta), _f32(np.asarray(x).T), _f32(y)
    n = xt.shape[1]
    grad = np.zeros(len(theta), np.float32)
    loss = ctypes.c_double(0.0)
    check(lib().gplum_mlp_loss_grad(abi.ptr(theta), abi.ptr(xt), abi.ptr(y), n, ctypes.byref(loss),
                                    abi.ptr(grad)))
    return loss.value, grad


def mlp_train(theta, x, y, n_steps, lr=3e-3):
    """Full-batch Adam on the device; returns (theta, loss_before, loss_after)."""
    theta, xt, y = _f32(theta).copy(), _f32(np.asarray(x).T), _f32(y)
    l0, l1 = ctypes.c_double(0.0), ctypes.c_double(0.0)
    check(lib().gplum_mlp_train(abi.ptr(theta), abi.ptr(xt), abi.ptr(y), xt.shape[1], n_steps, lr,
                                ctypes.byref(l0), ctypes.byref(l1)))
    return theta, l0.value, l1.value
