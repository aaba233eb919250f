"""CPU checks of the C ABI: libamm.so loads, exports every function that
include/amm.h declares, and rejects bad arguments before touching a GPU."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    src = (ROOT / "include" / "amm.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(amm_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    from eigenvaluecomputation_by_amm_amd import _lib
    assert sorted(_lib.exported_symbols()) == _declared()


def test_library_exports_every_declared_symbol():
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    for name in _declared():
        assert re.search(rf"\bT {name}\b", out), name


def test_abi_version():
    from eigenvaluecomputation_by_amm_amd import _lib
    assert _lib.load().amm_abi_version() == 1


@pytest.mark.parametrize("args,msg", [
    ((None, 0, 10, 10, 10), "A is NULL"),
    ((16, 7, 10, 10, 10), "unknown dtype"),
    ((16, 0, 0, 10, 10), "d and n"),
    ((16, 0, 10, 10, 8), "ld must be >= d"),
    ((16, 0, 10, 10, 10), "multiple of 16 bytes"),
    ((16, 1, 12, 10, 12), "multiple of 16 bytes"),
    ((18, 0, 8, 10, 8), "16-byte aligned"),
])
def test_ctx_create_validation(args, msg):
    """Argument checks run before any HIP call (safe without a GPU)."""
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    A, dt, d, n, ld = args
    rc = lib.amm_ctx_create(C.c_void_p(A) if A else None, dt, d, n, ld, None, C.byref(h))
    assert rc == _lib.AMM_E_INVALID if hasattr(_lib, "AMM_E_INVALID") else rc == -1
    assert msg in lib.amm_last_error().decode()
    assert not h.value


def test_null_context_calls_fail_cleanly():
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = _lib.load()
    assert lib.amm_ctx_destroy(None) == 0
    assert lib.amm_power_method(None, None, None, None, None) == -1
    assert "NULL" in lib.amm_last_error().decode()
