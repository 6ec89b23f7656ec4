"""The C-ABI libraries load and export every symbol declared in include/thcgw.h
(libthcgw.so) and include/thcgw_eri.h (libthcgw_eri.so); no GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(header="thcgw.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(thcgw_[a-z_0-9A-Z]+)\s*\(", txt)))


def test_header_matches_binding_list():
    from coqui_amd.thcgw import EXPORTS

    assert sorted(EXPORTS) == header_functions()


def test_library_exports_all_symbols():
    from coqui_amd.thcgw import LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.fail("libthcgw.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_binding_declares_signatures():
    from coqui_amd.thcgw import load_library

    lib = load_library()
    for f in header_functions():
        assert getattr(lib, f).argtypes is not None or f == "thcgw_destroy"


def test_last_error_without_context():
    """thcgw_last_error(NULL) is a pure host call: the empty create-error string."""
    from coqui_amd.thcgw import load_library

    lib = load_library()
    assert lib.thcgw_last_error(None) == b""


def test_eri_library_exports_all_symbols():
    from coqui_amd.eri import ERI_EXPORTS, ERI_LIB_PATH, load_eri_library

    if not os.path.exists(ERI_LIB_PATH):
        pytest.fail("libthcgw_eri.so not built: run __graft_entry__.build()")
    assert sorted(ERI_EXPORTS) == header_functions("thcgw_eri.h")
    lib = load_eri_library()
    for f in header_functions("thcgw_eri.h"):
        assert getattr(lib, f).argtypes is not None
