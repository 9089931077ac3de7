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


def test_front_end_exports_nvtt_api():
    """libnvtt_amd.so carries the nvtt:: entry points of the drop-in path."""
    import subprocess
    out = subprocess.run(["nm", "-DC", N.NVTT_LIB_PATH], capture_output=True, text=True, check=True).stdout
    for sym in ("nvtt::Compressor::compress(nvtt::Surface const&, int, int, nvtt::CompressionOptions const&, nvtt::OutputOptions const&) const",
                "nvtt::Compressor::enableCudaAcceleration(bool)",
                "nvtt::Surface::buildNextMipmap(nvtt::MipmapFilter, float, float const*, int)",
                "nvtt::Surface::setImage(nvtt::InputFormat, int, int, int, void const*)",
                "nvtt::Surface::toLinearFromSrgb()",
                "nvtt::OutputOptions::setOutputHandler(nvtt::OutputHandler*)",
                "nvtt::CompressionOptions::setFormat(nvtt::Format)"):
        assert sym in out, sym
    assert os.path.exists(os.path.join(N.LIB_DIR, "nvtt_api_test"))


def test_c_interface_exports():
    """Every function nvtt_wrapper.h declares is exported with C linkage, and
    the C driver built from that header reports the version without a GPU."""
    import re
    import subprocess
    hdr = open(os.path.join(N.PKG_DIR, "nvtt", "nvtt_wrapper.h")).read()
    names = set(re.findall(r"NVTT_WRAPPER_API\s+[^;(]*?\b(nvtt\w+)\s*\(", hdr))
    assert len(names) == 42
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(N.LIB_DIR, "libnvtt_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert names <= exported, sorted(names - exported)
    r = subprocess.run([os.path.join(N.LIB_DIR, "nvtt_wrapper_test"), "info"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and r.stdout.splitlines()[0] == "version 20102"
    assert len(r.stdout.splitlines()) == 8
