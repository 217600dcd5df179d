"""The C-ABI library loads without a GPU and exports every function include/dmsc_hip.h declares."""

import ctypes
import os
import re

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dmsc_hip.h")
LIB = os.path.join(ROOT, "dmsc_amd", "lib", "libdmsc_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(dmsc_[a-z_]+)\s*\(", text))
    return sorted(n for n in names if n not in {"dmsc_alloc_fn"})


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("dmsc_context_create", "dmsc_context_destroy", "dmsc_extract", "dmsc_extract_batch", "dmsc_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_module_imports():
    from dmsc_amd.csrc import dmsc_gpu

    for fn in ("extract_dmsc", "extract_dmsc_list", "profile_enable", "profile_read", "debug_fetch"):
        assert hasattr(dmsc_gpu, fn)
