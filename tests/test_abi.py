"""The C-ABI library builds, loads and exports every entry point declared in
include/flexmm_amd.h; argument errors come back as codes and messages (no
GPU work is started by these calls)."""
import ctypes
import os
import re

import pytest

from flexmm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "flexmm_amd.h")).read()
    return sorted(set(re.findall(r"\b(fmm2d_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from flexmm_amd import build
        build.build()
    return _lib.load()


def test_every_declared_symbol_is_exported_and_bound(lib):
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for %s" % n


def test_abi_version(lib):
    assert lib.fmm2d_abi_version() == _lib.ABI_VERSION


def test_operator_size(lib):
    for p, PD in ((60, 64), (32, 32)):
        assert lib.fmm2d_operator_size(p) == (50 * PD * PD + 2 * PD + 40 * PD + 20 * PD * PD
                                              + 64 + 40 * PD * PD)
    assert lib.fmm2d_operator_size(3) == -1
    assert lib.fmm2d_operator_size(65) == -1


def test_null_handles_are_argument_errors(lib):
    info = (ctypes.c_int64 * 16)()
    assert lib.fmm2d_tree_info(None, info, 16) == _lib.E_ARG
    assert b"null" in lib.fmm2d_last_error()
    assert lib.fmm2d_upward(None, None, 0, None) == _lib.E_ARG
    assert lib.fmm2d_plan_destroy(None) == 0
    assert lib.fmm2d_plan_create(0, None, 0, 60, 0, None) == _lib.E_ARG


def test_check_raises_with_message(lib):
    code = lib.fmm2d_downward(None, None)
    with pytest.raises(_lib.FMMError) as ei:
        _lib.check(code)
    assert ei.value.code == _lib.E_ARG
    assert "plan handle is null" in str(ei.value)
