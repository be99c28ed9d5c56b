"""CPU-side checks of the native library: it is built, it loads, it exports
every entry point include/odba/odba.h declares, and without a GPU the
context creation reports ODBA_E_NO_DEVICE instead of falling back."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "odba", "odba.h")
LIB = os.path.join(ROOT, "opendba_amd", "lib", "libodba_hip.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-j4", "-C", os.path.join(ROOT, "opendba_amd", "csrc")])
    return C.CDLL(LIB)


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|uint64_t|const char\*)\s+(odba_\w+)\(", txt, re.M)))


def test_header_declares_expected_api():
    syms = declared_symbols()
    for s in ["odba_ctx_create", "odba_dtw_all_pairs", "odba_dtw_path", "odba_dba_round", "odba_znorm"]:
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_lists_all_exports():
    from opendba_amd import native
    assert sorted(native.EXPORTS) == declared_symbols()


def test_no_device_is_an_error_not_a_fallback(lib):
    import torch  # noqa: F401  (the GPU box is where devices exist)
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    lib.odba_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    rc = lib.odba_ctx_create(0, C.byref(h))
    assert rc == 105  # ODBA_E_NO_DEVICE


def test_reference_style_caller_compiles_against_the_header(lib):
    """tests/cpp/drop_in_calls.cpp calls every template of
    include/odba/opendba.hpp with the reference's argument lists; it must
    compile and link unchanged (-Wall, no warnings)."""
    r = subprocess.run(["make", "-B", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr.lower(), r.stderr
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "_build", "drop_in_calls"))
