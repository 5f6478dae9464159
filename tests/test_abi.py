"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this suite runs without a GPU)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from pyfwi_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pyfwi_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_binding_exports():
    assert declared_functions() == set(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in sorted(declared_functions()):
        assert hasattr(lib, name), name


def test_version_and_error_string():
    lib = _native.load()
    assert lib.pyfwi_version() == 100
    assert isinstance(lib.pyfwi_last_error(), bytes)


def test_invalid_grid_is_rejected_before_any_device_call():
    lib = _native.load()
    g = _native.Grid()
    g.tnz, g.tnx, g.npml, g.sdo = 50, 50, 10, 3   # sdo must be 2 or 4
    g.dt, g.dx, g.dz = 1e-3, 5.0, 5.0
    h = ctypes.c_void_p()
    rc = lib.pyfwi_plan_create(ctypes.byref(g), None, ctypes.byref(h))
    assert rc == 1 and b"sdo" in lib.pyfwi_last_error()
    g.sdo, g.dt = 2, -1.0
    assert lib.pyfwi_plan_create(ctypes.byref(g), None, ctypes.byref(h)) == 1


def test_pointer_placement_is_checked():
    """The library takes caller addresses as they are; the binding refuses a
    host buffer passed as device memory (which a kernel would fault on) and a
    device tensor passed as host memory."""
    import torch
    from pyfwi_amd import _native
    a = np.zeros(4, np.float32)
    assert _native.ptr(a, _native.MEM_HOST) == a.ctypes.data
    with pytest.raises(ValueError):
        _native.ptr(a, _native.MEM_DEVICE)
    t = torch.zeros(4)
    assert _native.ptr(t, _native.MEM_HOST) == t.data_ptr()
    with pytest.raises(ValueError):
        _native.ptr(t, _native.MEM_DEVICE)
    with pytest.raises(ValueError):
        _native.ptr(torch.zeros(4, 2).t(), _native.MEM_HOST)
