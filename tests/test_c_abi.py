"""The C ABI: every entry point declared in include/casadi_hip.h is exported
by the in-tree library, and host-only calls behave without a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "casadi_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(cx[hs]_[A-Za-z0-9_]+)\s*\(", src, re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("cxh_tape_create", "cxh_map_eval", "cxh_map_fwd_adj", "cxh_map_fwd_ckpt", "cxh_map_rev_ckpt",
                 "cxh_jac_create", "cxh_map_jac", "cxh_batch_create", "cxh_batch_launch", "cxh_batch_tune",
                 "cxh_ldl_create", "cxh_ldl_factor_solve", "cxh_qr_create", "cxh_qr_factor_solve",
                 "cxh_tape_engine", "cxs_function_new", "cxs_sp_ldl"):
        assert must in names, must


def test_library_exports_every_declared_symbol():
    from casadi_amd import _lib as L
    missing = [n for n in declared_functions() if not hasattr(L.lib, n)]
    assert not missing, missing


def test_errors_cross_as_return_codes():
    from casadi_amd import _lib as L
    rc = L.cxh_tape_create(None, -1, 1, 0, None, 0, None, C.byref(L.P()))
    assert rc < 0 and b"negative" in L.cxh_last_error()
    assert L.cxh_tape_engine(None, 1) < 0
    assert L.cxh_tape_config(None, 64, 1, 64) < 0
