"""Structural checks on the optimised device IR of the render kernels (CPU only).

Both persistent kernels pop work from a global counter with one lane and
broadcast the index with readfirstlane.  That is only correct if every lane
of the wave reaches the readfirstlane together.  An optimiser may thread a
branch on the lane index at the end of the loop body into the lane-0 fetch
at its head, so that lanes 1-63 loop back to the readfirstlane without lane 0
and read a stale index forever.  The IR must show, for every kernel
instantiation, that the broadcast value is a phi of exactly the atomic's
result and the constant 0 of the lanes that skipped it.  (The cooperative
kernel has two such pops: the fetch of a ray when none is claimed, and the
claim of the next one while the current ray is traced.)
"""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "cuda-csg-tree-raycasting_amd")


@pytest.fixture(scope="module")
def device_ir():
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    subprocess.run(["make", "-s", "-C", PKG, "ir"], check=True, capture_output=True, timeout=600)
    with open(os.path.join(PKG, "_build", "render.ll")) as f:
        return f.read()


def _kernels(ir):
    return re.findall(r"^define[^\n]*@(_Z\w*(?:Coop|Render)Kernel\w*)\([^\n]*\{\n(.*?)\n\}", ir, re.S | re.M)


def test_every_kernel_instantiation_is_present(device_ir):
    names = [n for n, _ in _kernels(device_ir)]
    # fast x count x ((LDS | global | global with the top of the tree in LDS) x XCD homes | batch)
    assert sum("RenderKernel" in n for n in names) == 28
    assert sum("CoopKernel" in n for n in names) == 8      # fast x count x batch


def test_queue_pop_broadcast_is_not_threaded(device_ir):
    for name, body in _kernels(device_ir):
        defs = dict(re.findall(r"^\s*(%\d+) = (.*)$", body, re.M))
        pops = 0
        for arg in re.findall(r"readfirstlane\.i32\(i32 (%\d+)\)", body):
            d = defs.get(arg, "")
            if not d.startswith("phi"):
                continue
            ins = re.findall(r"\[ ([^,]+), (%\d+) \]", d)
            srcs = [defs.get(v, v) for v, _ in ins]
            if not any(s.startswith("atomicrmw add") for s in srcs):
                continue
            pops += 1
            assert len(ins) == 2 and "0" in [v for v, _ in ins], (name, d)
        assert pops == (2 if "CoopKernel" in name else 1), name
