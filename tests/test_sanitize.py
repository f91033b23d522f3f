"""The host side of the C ABI under AddressSanitizer and UndefinedBehavior-
Sanitizer (no GPU): the GPU-free ABI and host-logic tests run in a child
process against lib/sanitize/libdfp.so (make -C dopamine-fdm_amd/csrc
sanitize: the same sources with -fsanitize=address,undefined on the host
pass) with the clang ASan runtime preloaded; any report fails the run."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "dopamine-fdm_amd", "lib", "sanitize", "libdfp.so")
ASAN_RT = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))

CHILD = r"""
import os, sys
sys.path[:0] = [{root!r}, os.path.join({root!r}, "dopamine-fdm_amd"), os.path.join({root!r}, "tests")]
from dfp import _lib
assert _lib.LIB_PATH == {lib!r}, _lib.LIB_PATH
_lib.lib()
maps = open("/proc/self/maps").read()
assert "libclang_rt.asan" in maps and {lib!r} in maps, "sanitized library or ASan runtime not loaded"
import pytest
sys.exit(pytest.main(["-q", "-p", "no:cacheprovider", "-m", "not gpu",
                      os.path.join({root!r}, "tests", "test_abi.py"),
                      os.path.join({root!r}, "tests", "test_host_logic.py")]))
"""


@pytest.mark.skipif(not os.path.exists(SAN_LIB) or not ASAN_RT,
                    reason="sanitizer build absent (make -C dopamine-fdm_amd/csrc sanitize) or no clang ASan runtime")
def test_abi_and_host_logic_under_asan_and_ubsan():
    env = dict(os.environ)
    env.update(DFP_LIB=SAN_LIB, LD_PRELOAD=ASAN_RT[-1],
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=87")
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT, lib=SAN_LIB)], env=env, cwd="/tmp",
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
