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
    # the batched tuners and timers check their handles before any HIP call
    assert lib.amm_batch_autotune_run(None, None, 1, None) == -1
    assert "NULL" in lib.amm_last_error().decode()
    assert lib.amm_batch_time_dense(None, 1, None) == -1
    assert lib.amm_batch_autotune(None, 1, None) == -1
    assert lib.amm_batch_destroy(None) == 0


def test_header_compiles_as_c_and_links(tmp_path):
    """include/amm.h is plain C (C99): a C program that calls the library's
    host-only entry points compiles with gcc, links against libamm.so and
    runs without a GPU (argument validation happens before any device call)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = Path(_lib.load()._name)
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include <string.h>\n#include "amm.h"\n'
        "int main(void) {\n"
        "  amm_ctx *ctx = 0;\n"
        "  if (amm_abi_version() != 1) return 1;\n"
        "  if (amm_ctx_create(0, AMM_DTYPE_F32, 10, 10, 10, 0, &ctx) == 0) return 2;\n"
        '  if (!strstr(amm_last_error(), "NULL")) return 3;\n'
        '  printf("ok %s\\n", amm_last_error());\n'
        "  return 0;\n}\n")
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                    f"-L{lib.parent}", "-lamm", f"-Wl,-rpath,{lib.parent}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.startswith("ok")


def test_noalloc_exact_product_rejects_before_any_hip_call():
    """amm_exact_product_noalloc validates its arguments and sizes its plan
    against the fixed scratch before it touches the device."""
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = _lib.load()
    p = C.c_void_p(16)
    assert lib.amm_exact_product_noalloc(None, 0, 8, 8, 8, p, p, None) == -1
    assert "A is NULL" in lib.amm_last_error().decode()
    assert lib.amm_exact_product_noalloc(p, 0, 8, 8, 8, None, p, None) == -1
    assert lib.amm_exact_product_noalloc(p, 0, 10, 8, 10, p, p, None) == -1
    assert "multiple of 16 bytes" in lib.amm_last_error().decode()
    # 2^21 rows: the padded x and y alone fill the 16 MiB scratch
    assert lib.amm_exact_product_noalloc(p, 0, 1 << 21, 4, 1 << 21, p, p, None) == -1
    assert "scratch" in lib.amm_last_error().decode()
    # 2^20 columns of 2^17 rows: even chunks of 2048 columns (the largest) leave
    # 512 partials of A t at 512 KiB each, far beyond the scratch
    assert lib.amm_exact_product_noalloc(p, 0, 1 << 17, 1 << 20, 1 << 17, p, p, None) == -1
    assert "scratch" in lib.amm_last_error().decode()
