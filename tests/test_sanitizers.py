"""Host code under the sanitizers (CPU only): the oracle's pipeline and the
library's host Delaunay step (worker pool, split merge, block cache,
namespaced Triangle copies) built with AddressSanitizer + UndefinedBehavior
Sanitizer, and again with ThreadSanitizer, by oracle/build_sanitized.py;
sanitize_main.cpp checks every mesh against the oracle's serial Triangle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

ENV = {
    "asan": {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=66", "UBSAN_OPTIONS": "halt_on_error=1"},
    "tsan": {"TSAN_OPTIONS": "halt_on_error=1:exitcode=66"},
}


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_is_sanitizer_clean(kind):
    if not os.path.isfile("/root/reference/src/triangle.cpp"):
        pytest.skip("the sanitized build compiles Triangle from the reference checkout")
    import sys
    sys.path.insert(0, ROOT)
    from oracle import build_sanitized
    exe = build_sanitized.build(kind)
    env = dict(os.environ)
    env.update(ENV[kind])
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-4000:]
    assert "Sanitizer" not in text and "runtime error" not in text, text[-4000:]
    assert "0 mesh mismatches, 0 pipeline failures" in r.stdout
