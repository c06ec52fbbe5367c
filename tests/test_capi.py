"""CPU checks of the C ABI: the library loads and exports every symbol include/*.h declares,
and the struct layouts match the reference's byte sizes.  No compute calls (no GPU here)."""
import ctypes

from gplum_amd import abi, gpu


def test_layouts():
    abi.check_layouts()


def test_library_exports_declared_symbols():
    L = gpu.lib()
    names = gpu.declared_symbols()
    assert "gplum_calc_force_epep" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_params_default_matches_reference():
    p = gpu.params()
    # FPGrav static initialisers, src/particle.h:917-961
    assert (p.R_cut0, p.R_cut1, p.R_search0, p.R_search1) == (2.0, 8.0, 1.0, 4.0)
    assert p.dt_tree == 2.0 ** -5 and p.dt_min == 2.0 ** -13
    assert p.gamma == 0.5 and p.g_1_inv == -2.0 and p.g_1_inv7 == -128.0
    assert p.n_leaf_limit == 8 and p.n_group_limit == 256
