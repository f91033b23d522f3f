"""Every HIP kernel of the library keeps its working set in registers: a
kernel with private (scratch) memory per lane spills register arrays to
HBM-backed memory, which is a silent several-fold slowdown of the LDS
transforms (their butterfly outputs live in register arrays between the read
and write phases).  Compiles each translation unit for gfx950 with the
compiler's kernel resource report and checks ScratchSize of every kernel.

Known exception: the persistent (batched) FP64 forward x kernels, which read
their case record from memory instead of kernel arguments, spill a few
values (12-44 bytes per lane); they are held to that bound."""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dopamine-fdm_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "--cuda-device-only", "-Rpass-analysis=kernel-resource-usage"]


def _report(src, out_dir):
    obj = os.path.join(out_dir, os.path.basename(src) + ".o")
    r = subprocess.run([HIPCC, *FLAGS, "-c", src, "-o", obj], capture_output=True, text=True, cwd=CSRC)
    assert r.returncode == 0, r.stderr[-2000:]
    kernels, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            kernels[name] = int(m.group(1))
    return kernels


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="no hipcc")
def test_no_kernel_uses_scratch():
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    out_dir = tempfile.mkdtemp()
    try:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            reports = list(ex.map(lambda s: _report(s, out_dir), srcs))
    finally:
        shutil.rmtree(out_dir, ignore_errors=True)
    kernels = {k: v for rep in reports for k, v in rep.items()}
    assert len(kernels) > 50
    batched_fwd = re.compile(r"^_ZN3dfp13kb_div_rfft_xILi[01]E15HIP_vector_typeIdLj2EE")
    spills = {k: v for k, v in kernels.items() if v and not (batched_fwd.match(k) and v <= 48)}
    assert not spills, spills
