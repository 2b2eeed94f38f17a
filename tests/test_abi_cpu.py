"""The C-ABI library loads on the CPU and exports every entry point that
include/ymrt.h declares (no GPU calls are made)."""
import ctypes
import os
import re

import ymrt_amd

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ymrt.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ymrt_[a-z_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(ymrt_amd.EXPORTS)


def test_library_exports_every_symbol():
    L = ctypes.CDLL(ymrt_amd.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_abi_version_and_error_string():
    L = ymrt_amd.lib()
    assert L.ymrt_abi_version() == 2
    assert isinstance(L.ymrt_last_error(), bytes)
