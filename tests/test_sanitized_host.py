"""The oracle's C restatement built with AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/sanitize.mk: gcc, host code only) and
driven through its entry points, on odd and one-texel sizes, in a child
process with the sanitizer runtime preloaded: it must run to the end with no
sanitizer report."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

SCRIPT = r"""
import sys
import numpy as np
sys.path[:0] = [sys.argv[1], sys.argv[2]]
import oracle_lib
from ntx_amd import synth
o = oracle_lib.Oracle(sys.argv[3])
rng = np.random.default_rng(3)
for (w, h) in ((1, 1), (3, 5), (7, 1), (13, 9), (37, 23)):
    img = synth.to_planar(synth.noise_rect_rgba(w, h, seed=w * h, transparent=True)).reshape(-1)
    noisy = (rng.random(4 * w * h, dtype=np.float32) * np.float32(1.4) - np.float32(0.2))
    for src in (img, noisy):
        for fmt in (1, 3, 4, 5, 6):
            for q in (0, 1, 3):
                for am in (0, 1):
                    try:
                        o.compress_q(fmt, q, src, w, h, alpha_mode=am, threads=2)
                    except Exception:
                        pass
        for filt in (0, 1, 2):
            for wrap in (0, 1, 2):
                if (w, h) != (1, 1):  # no next level (Surface.cpp:155-167)
                    o.build_next_mipmap(filt, (0.5, 1.0, 3.0)[filt], None if filt < 2 else (4.0, 1.0), wrap, src, w, h)
        o.srgb_to_linear(src.copy())
blk = np.random.default_rng(5).integers(0, 256, (64, 16, 4)).astype(np.uint8)
for b in blk:
    o.bc1a_block(b, 1); o.bc3n_block(b, 0); o.bc3_color_block(b, 1); o.bc4_channel(b[:, 0].copy())
print("sanitized oracle ok")
"""


def test_oracle_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-f", "sanitize.mk", "liboracle_san.so"], cwd=ORACLE, capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed: " + r.stderr[-400:])
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan):
        pytest.skip("libasan not found")
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", SCRIPT, os.path.join(ROOT, "tests"),
                        os.path.join(ROOT, "nvidia-texture-tools_amd"), os.path.join(ORACLE, "liboracle_san.so")],
                       capture_output=True, text=True, timeout=900, env=env)
    report = r.stderr[-2000:]
    assert r.returncode == 0 and "sanitized oracle ok" in r.stdout, report
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, report
