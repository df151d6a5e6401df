"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/dompap_hip.h declares, the ctypes signatures cover them, and the
hiprtc pair kernels compile for gfx950."""
import ctypes
import os
import re

import pytest

from dompap_amd import _lib, codegen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dompap_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dp_\w+)\(", text, flags=re.M)))


def test_header_symbols_exported():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name
    assert lib.dp_version() == 100


def test_errors_without_device_are_reported():
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    rc = lib.dp_create(ctypes.byref(ctx), 0, 3, 10, 1, 0)
    if rc == 0:  # a device is present: clean up
        lib.dp_destroy(ctx)
        return
    assert rc < 0 and lib.dp_last_error(None)
    assert lib.dp_create(ctypes.byref(ctx), 0, 9, 10, 1, 0) == _lib.DP_ERR_ARG
    assert lib.dp_create(ctypes.byref(ctx), 0, 5, 10, 1, 0) == _lib.DP_ERR_ARG


@pytest.mark.parametrize("pot", ["4*(r**-12-r**-6)", "(1-r)**2", "Harmonic repulsive", "r**-12",
                                 "(((1-r)**2)**(1/2))**(7/2)", "exp(-r)/r"])
def test_pair_kernels_compile(pot):
    lib = _lib.load()
    code = codegen.generate(pot)
    for dim in (1, 2, 3, 4):
        for mode in (0, 1):
            log = ctypes.create_string_buffer(1 << 15)
            size = ctypes.c_int64()
            rc = lib.dp_compile_check(dim, mode, 2, int(code.even), code.source.encode(), 0, log, 1 << 15,
                                      ctypes.byref(size))
            assert rc == 0, log.value.decode()
            assert size.value > 1000
