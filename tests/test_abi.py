"""The C ABI library loads and exports every entry point include/ntx_hip.h
declares; host-only behaviour that needs no device."""
import ctypes
import os
import re

import ntx_amd as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ntx_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"NTX_API\s+[\w\s\*]+?\b(ntx_hip_\w+)\s*\(", src)))


def test_header_declarations_match_binding():
    assert declared_symbols() == sorted(N.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing


def test_compressed_size_matches_reference_formula():
    lib = N.load_library()
    for fmt, bs in ((N.FORMAT_BC1, 8), (N.FORMAT_BC3, 16), (N.FORMAT_BC4, 8), (N.FORMAT_BC5, 16)):
        for w, h in ((1, 1), (4, 4), (5, 3), (1024, 1024), (4097, 2)):
            assert lib.ntx_hip_compressed_size(fmt, w, h) == ((w + 3) // 4) * ((h + 3) // 4) * bs
            assert N.compressed_size(fmt, w, h) == lib.ntx_hip_compressed_size(fmt, w, h)


def test_status_strings():
    lib = N.load_library()
    for s in range(6):
        assert lib.ntx_hip_status_string(s)


def test_null_arguments_rejected_without_device():
    lib = N.load_library()
    assert lib.ntx_hip_device_count(None) == 1
    assert lib.ntx_hip_create(0, None) == 1
    assert lib.ntx_hip_destroy(None) == 0
