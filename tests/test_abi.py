"""The C-ABI library loads and exports every entry point include/cmi.h
declares; host-only entry points agree with the oracle.  No device needed."""
import ctypes as C
import os
import re

import pytest

import oracle_lib as O
import parallelcmacionize_amd as cmi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cmi_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = cmi.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n
    # and the Python front end binds exactly those
    assert set(cmi.SIGNATURES) == names


def test_host_direction_algebra_matches_oracle(oracle):
    for d in range(27):
        assert cmi.output_to_input(d) == oracle.orc_output_to_input_direction(d)
    assert cmi.output_to_input(27) == -1 and cmi.output_to_input(-1) == -1
    for mask in range(64):
        assert cmi.direction_from_mask(mask) == oracle.orc_direction_from_mask(mask)
    assert cmi.direction_from_mask(64) == -1


def test_create_rejects_bad_arguments_without_device():
    lib = cmi.load_library()
    import numpy as np
    h = C.c_void_p()
    box = np.array([0., 0., 0., 1., 1., 1.])
    bad = np.array([10, 10, 10], np.int32)
    nsub = np.array([3, 3, 3], np.int32)  # 10 % 3 != 0
    rc = lib.cmi_create(box.ctypes.data_as(C.POINTER(C.c_double)),
                        bad.ctypes.data_as(C.POINTER(C.c_int)),
                        nsub.ctypes.data_as(C.POINTER(C.c_int)), 0,
                        C.byref(h))
    assert rc == -1 and not h.value
    assert lib.cmi_last_error(None) == b"null context"


def test_no_cpu_fallback_in_product_package():
    """The product package never loads the oracle."""
    pkg = os.path.join(ROOT, "parallelcmacionize_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "liboracle" not in text and "oracle_lib" not in text, fn
                assert "libref" not in text, fn
