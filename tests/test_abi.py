"""The C-ABI library loads and exports every entry point include/gev_pm.h
declares (no GPU calls are made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gev_pm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(gev_[a-z_0-9A-Z]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for must in ["gev_projection_T00_project", "gev_projection_Tij_project", "gev_projection_T0i_project",
                 "gev_prepare_ft_source_scalar", "gev_prepare_ft_source_tensor", "gev_fft_execute",
                 "gev_solve_modified_poisson_ft", "gev_project_ft_scalar", "gev_evolve_ft_vector",
                 "gev_project_ft_vector", "gev_update_vel", "gev_move_particles", "gev_run_cycles"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import gevpm
    lib = ctypes.CDLL(gevpm.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding knows the same list
    assert sorted(gevpm.EXPORTED) == declared_functions()


def test_host_background_matches_oracle(orc):
    """gev_hconf / gev_rungekutta4bg / gev_next_dtau are host code: no GPU
    needed, and they must equal the oracle bit for bit (same expressions)."""
    import gevpm
    cg = gevpm.make_cosmology()
    co = orc.make_cosmo()
    for a in [0.01, 0.0099009900990099, 0.3, 1.0]:
        for fpg in [0.0171, 0.7]:
            assert gevpm.hconf(a, fpg, cg) == orc.hconf(a, fpg, co)
            assert gevpm.rungekutta4bg(a, fpg, cg, 0.01) == orc.rk4bg(a, fpg, co, 0.01)


def test_footprint_of_both_forms():
    """gev_ctx_footprint_bytes is host arithmetic (no GPU): the FP32 form
    halves the field and k-space storage and keeps the FP64 particles."""
    import gevpm
    N, cap = 512, 512 ** 3
    f64 = gevpm.footprint_bytes(N, cap, gevpm.FORM_FP64)
    f32 = gevpm.footprint_bytes(N, cap, gevpm.FORM_FP32)
    fields64 = 12 * N ** 3 * 8 + 11 * (N // 2 + 1) * N * N * 16
    particles = cap * (2 * (6 * 8 + 8) + 2 * 4)
    assert f64 >= fields64 + particles
    assert f64 - f32 >= fields64 // 2 - N * 8  # every field element halves
    assert f32 >= fields64 // 2 + particles
    assert gevpm.footprint_bytes(N, cap, 5) < 0  # unknown form


def test_bandwidth_probe_rejects_bad_arguments_before_any_device_work():
    import ctypes
    import gevpm
    L = gevpm.lib()
    g, m = ctypes.c_double(), ctypes.c_double()
    for args in ((0, 0, 1), (1 << 30, 7, 1), (1 << 30, 0, 0), (1 << 30, -1, 5)):
        assert L.gev_bandwidth_probe(*args, ctypes.byref(g), ctypes.byref(m)) == -1
