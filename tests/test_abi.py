"""The C ABI: libpdcs_hip.so loads and exports every symbol include/pdcs.h declares.

No compute call is made here (this runs without a GPU); only host-only entry
points and the error path of context creation are exercised.
"""
import ctypes as C
import os
import re

from pdcs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pdcs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pdcs_[a-z_]+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_library_exports_all_declared():
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_params_default_matches_reference_defaults():
    p = _lib.default_params()
    assert p.max_outer_iter == 10000 and p.max_inner_iter == 500000
    assert p.abs_tol == 1e-6 and p.rel_tol == 1e-6
    assert p.check_terminate_freq == 2000 and p.duality_gap_restart_freq == 2000
    assert p.use_duality_gap_restart == 1 and p.use_kkt_restart == 0
    assert p.use_preconditioner == 1


def test_struct_sizes_match_header_layout():
    # pdcs_params / pdcs_result / pdcs_cones as laid out by a C compiler
    assert C.sizeof(_lib.Cones) == 8 * 8
    assert C.sizeof(_lib.Params) == 8 * 5 + 8 * 3 + 4 * 10 + 8 * 3 + 4 * 2 + 8 * 4 + 8
    assert C.sizeof(_lib.Result) == 4 * 2 + 8 * 3 + 8 * 17 + 4 * 4 + 8 * 2


def test_ctx_create_without_device_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        return
    L = _lib.load()
    h = C.c_void_p()
    rc = L.pdcs_ctx_create(0, C.byref(h))
    assert rc != 0 and not h
    assert L.pdcs_last_error(None)
