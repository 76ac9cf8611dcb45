"""Host sanitizer runs (CPU only): the C oracle and the host side of the C-ABI
built with AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize/Makefile) and run to completion with both set to abort on
the first finding.  The oracle driver runs every oracle function on seeded
scenes with exactly sized heap arrays; the C-ABI driver, generated from
include/lig_abi.h, calls all of its entry points with zero, small, negative
and placeholder arguments (placeholder device pointers only where no GPU is
visible, so every call stops at its first HIP call)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _build(target):
    res = subprocess.run(["make", "-C", SAN, target], capture_output=True, text=True,
                         timeout=900)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]


def _run(exe):
    res = subprocess.run([os.path.join(SAN, "build", exe)], capture_output=True, text=True,
                         timeout=600, env=ENV)
    out = res.stdout + res.stderr
    assert res.returncode == 0 and "ok" in res.stdout, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    return res.stdout


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_under_address_and_ub_sanitizers():
    _build("oracle")
    out = _run("oracle_asan")
    assert out.count("scene ") == 26 and out.count("upscale ") == 4


@pytest.mark.skipif(not os.path.exists(os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")),
                    reason="needs hipcc")
def test_abi_host_side_under_address_and_ub_sanitizers():
    _build("abi")
    out = _run("abi_asan")
    assert "ok" in out
