"""The C-ABI library loads and exports every entry point include/fbc_gpu.h
declares (no device calls: these run without a GPU)."""
import ctypes
import os
import re

import fastbc_amd as F
from conftest import ROOT


def declared_functions(header="fbc_gpu.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fbc_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(F.EXPORTS)
    assert declared_functions("fbc_host.h") == sorted(F.HOST_EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(F.LIB_PATH)
    for name in declared_functions() + declared_functions("fbc_host.h"):
        assert hasattr(lib, name), name


def test_null_arguments_rejected_without_device():
    lib = F.lib()
    assert lib.fbc_create(0, None) == -1
    assert lib.fbc_set_tuning(None, 0, 0) == -1
    assert lib.fbc_last_stats(None, None, None) == -1
    assert lib.fbc_last_error(None) == b"no context"


def test_option_names_documented():
    """Every name fbc_set_option accepts is documented in include/fbc_gpu.h
    and in the Python wrapper's set_option."""
    src = open(os.path.join(ROOT, "fast-bc_amd", "csrc", "fbc_capi.cpp")).read()
    names = set(re.findall(r'strcmp\(name, "([a-z_0-9]+)"\)', src))
    assert {"delta", "autotune", "lds_view", "tune_log"} <= names
    header = open(os.path.join(ROOT, "include", "fbc_gpu.h")).read()
    doc = F.Context.set_option.__doc__
    for n in sorted(names):
        assert f'"{n}"' in header, n
        assert f'"{n}"' in doc, n
