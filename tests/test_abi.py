"""The C-ABI library loads, exports every declared entry point and rejects bad
arguments with a status (no compute calls: these run without a GPU)."""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

from spac_amd import abi
from spac_amd.fk import problem_desc

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def test_headers_exist():
    names = [h.name for h in HEADERS]
    assert "spac_fk.h" in names


@pytest.mark.parametrize("header", HEADERS, ids=lambda p: p.name)
def test_library_exports_every_declared_symbol(header):
    lib = abi.load()
    declared = abi.declared_symbols(header)
    assert declared, header
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    missing = [n for n in declared if n not in exported]
    assert not missing, missing


def test_abi_version_and_error_channel():
    lib = abi.load()
    assert lib.spac_abi_version() == 3
    assert lib.spac_fk_problem_create(None, None) == abi.SPAC_E_INVALID_ARG
    assert "null" in abi.last_error(lib)


def test_unregistered_plugin_rejected_before_device():
    lib = abi.load()
    d = problem_desc(coef=99)
    h = C.c_void_p()
    assert lib.spac_fk_problem_create(C.byref(d), C.byref(h)) == abi.SPAC_E_UNREGISTERED_PLUGIN
    assert "no device implementation" in abi.last_error(lib)
    d = problem_desc(domain=(1.0, 0.0, 0.0, 1.0))
    assert lib.spac_fk_problem_create(C.byref(d), C.byref(h)) == abi.SPAC_E_INVALID_ARG


def test_no_device_is_a_status_not_a_crash():
    lib = abi.load()
    if lib.spac_device_count() > 0:
        pytest.skip("a device is visible here")
    d = problem_desc()
    h = C.c_void_p()
    assert lib.spac_fk_problem_create(C.byref(d), C.byref(h)) == abi.SPAC_E_NO_DEVICE


def test_null_handle_calls_fail_cleanly():
    lib = abi.load()
    stats = (abi.PointStats * 1)()
    x = (C.c_double * 1)(0.0)
    assert lib.spac_fk_solve_points(None, x, x, 1, 10, 1e-3, 0, stats) == abi.SPAC_E_INVALID_ARG
    assert lib.spac_fk_simulate(None, None, 0, None, None) == abi.SPAC_E_INVALID_ARG
    assert lib.spac_fk_problem_destroy(None) == abi.SPAC_OK
