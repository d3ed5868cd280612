"""The C-ABI library loads on the CPU host and exports every declared symbol."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)  # declarations only
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+[\s*]+(pycwb_\w+)\s*\(", txt,
                                    flags=re.M))
    return names


def test_library_exports_declared_symbols():
    from pycwb_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared but not exported"
    assert set(_lib.SIGNATURES) <= declared


def test_workspace_query_is_pure_host():
    from pycwb_amd import _lib

    lib = _lib.load()
    n = lib.pycwb_sky_scan_workspace_bytes(3, 1024, 129)
    assert n == 256 + 3 * 1024 * 129 * 8
    assert b"gfx950" in lib.pycwb_version()


def test_batch_create_rejects_bad_desc():
    from pycwb_amd import _lib

    lib = _lib.load()
    d = _lib.BatchDesc()
    out = ctypes.c_void_p()
    assert lib.pycwb_batch_create(ctypes.byref(d), ctypes.byref(out)) == _lib.PYCWB_ERR_ARG
    assert out.value is None
