"""The C ABI's host code under AddressSanitizer, on the CPU.

`make asan` builds the library with its host pass instrumented; the C-ABI
checks of tests/test_cabi_cpu.py (exports, host helpers such as the
float64 rain interpolation, argument validation of every entry point with
null, undersized or inconsistent arguments, the no-device error path) then
run in a child process that loads that build (SWE_LIB_PATH) with the clang
ASan runtime preloaded.  Any heap, stack or global overflow, use after free
or double free in those paths aborts the child.
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "gpu-hybrid-flood-forecasting_amd")


def _asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def test_c_abi_checks_under_address_sanitizer():
    runtime = _asan_runtime()
    if runtime is None:
        pytest.skip("the clang AddressSanitizer runtime is not installed")
    r = subprocess.run(["make", "-s", "-j6", "-C", PKG, "asan"], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lib = os.path.join(PKG, "hybrid_flood_amd", "libswe_hip_asan.so")
    env = dict(os.environ, SWE_LIB_PATH=lib, LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1")
    run = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                          "tests/test_cabi_cpu.py", "-m", "not gpu"],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = run.stdout[-3000:] + run.stderr[-3000:]
    assert run.returncode == 0, out
    assert "AddressSanitizer" not in run.stderr, out
    assert " passed" in run.stdout, out
