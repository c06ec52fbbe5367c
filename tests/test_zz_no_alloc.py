"""The allocation-free EP-EP entry point (gplum_calc_force_epep_noalloc) against the
allocating one (gplum_calc_force_epep) on the same inputs: bit-identical Force arrays on
two calls in a row, the second with fewer i- and j-particles (the per-device chunk table
keeps the first call's entries beyond the second's and must not be read), accumulation
into a non-cleared Force array, and the size limit reported as an argument error.

Device buffers for the calls come from the HIP runtime the library itself links
(hipMalloc through ctypes in the test, not in the entry point)."""
import ctypes

import numpy as np
import pytest

from gplum_amd import abi, gpu

pytestmark = pytest.mark.gpu

NOALLOC_MAX_NI = 4096 * 256


def make_lists(rng, ni, nj):
    """An i-group and a j-list around it: neighbours at a few r_search, far pairs too."""
    epi = np.zeros(ni, dtype=abi.EPI)
    epi["id_local"] = np.arange(ni)
    epi["pos"] = np.array([1.0, 0.2, 0.01]) + rng.normal(0, 3e-4, (ni, 3))
    epi["r_out"] = 1e-4 * (1 + rng.uniform(0, 0.5, ni))
    epi["r_search"] = 2.5 * epi["r_out"]
    epj = np.zeros(nj, dtype=abi.EPJ)
    epj["id_local"] = np.arange(nj)            # the first ni j-particles are the i-particles
    epj["id"] = np.arange(nj)
    k = min(ni, nj)
    epj["pos"][:k] = epi["pos"][:k]
    epj["r_out"][:k] = epi["r_out"][:k]
    epj["r_search"][:k] = epi["r_search"][:k]
    epj["pos"][k:] = np.array([1.0, 0.2, 0.01]) + rng.normal(0, 1e-3, (nj - k, 3))
    epj["r_out"][k:] = 1e-4
    epj["r_search"][k:] = 2.5e-4
    epj["mass"] = rng.uniform(0.5e-12, 2e-12, nj)
    return epi, epj


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libgplum_gpu.so is linked with."""
    H2D, D2H = 1, 2

    def __init__(self):
        gpu.lib()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, arr):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(arr.nbytes, 256)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, abi.ptr(arr), arr.nbytes, self.H2D) == 0
        return p

    def download(self, p, arr):
        assert self.L.hipMemcpy(abi.ptr(arr), p, arr.nbytes, self.D2H) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def noalloc_call(hip, epi, epj, force, eps2):
    d_epi, d_epj, d_f = hip.upload(epi), hip.upload(epj), hip.upload(force)
    gpu.check(gpu.lib().gplum_calc_force_epep_noalloc(d_epi, len(epi), d_epj, len(epj), d_f, eps2, None))
    hip.sync()
    out = np.empty_like(force)
    hip.download(d_f, out)
    hip.free()
    return out


def test_noalloc_epep_matches_allocating_entry_point():
    hip = Hip()
    rng = np.random.default_rng(11)
    eps2 = np.float32(1e-12)
    for ni, nj in ((1000, 2600), (300, 700)):
        print("no-alloc EP-EP: ni=%d nj=%d" % (ni, nj), flush=True)
        epi, epj = make_lists(rng, ni, nj)
        start = abi.cleared_force(ni)
        ref = start.copy()
        gpu.calc_force_epep(epi, epj, ref, eps2)
        got = noalloc_call(hip, epi, epj, start, eps2)
        assert got.tobytes() == ref.tobytes()
        assert (got["neighbor"]["number"] > 0).any()
        # accumulation into a non-cleared Force array, as the functor does
        ref2 = ref.copy()
        gpu.calc_force_epep(epi, epj, ref2, eps2)
        got2 = noalloc_call(hip, epi, epj, ref, eps2)
        assert got2.tobytes() == ref2.tobytes()


def test_noalloc_limit_is_an_argument_error():
    hip = Hip()
    p = hip.upload(np.zeros(64, dtype=np.uint8))
    rc = gpu.lib().gplum_calc_force_epep_noalloc(p, NOALLOC_MAX_NI + 1, p, 1, p, 0.0, None)
    hip.free()
    assert rc != 0
    assert b"no-allocation limit" in gpu.lib().gplum_last_error()
