"""Register and scratch budget of the hot kernels, from hipcc's resource report
for gfx950 (no GPU needed).

Every k_resident instance has to fit 96 VGPRs (five waves per SIMD) with no
scratch: a single value spilled to scratch costs the batched workload about 3 %,
and the register allocation sits at its limit, so small source changes can
tip it.  k_strip has to stay free of scratch as well.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"]


def resources(source, tmp_path):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    out = subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", *FLAGS, "-c", "-o",
                          str(tmp_path / "k.o"), os.path.join(ROOT, "mcrg_amd", "csrc", source)],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    kernels, name = {}, None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels[name] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and name:
            kernels[name][m.group(1).split()[0]] = int(m.group(2))
    return kernels


def test_resident_kernels_fit_without_scratch(tmp_path):
    ks = {n: r for n, r in resources("resident_kernels.hip", tmp_path).items() if "k_resident" in n}
    assert len(ks) == 5, sorted(ks)  # workgroups of 64 .. 1024 threads
    for name, r in ks.items():
        assert r["ScratchSize"] == 0, (name, r)
        assert r["VGPRs"] <= 96 and r["Occupancy"] >= 5, (name, r)


def test_strip_kernels_without_scratch(tmp_path):
    ks = {n: r for n, r in resources("strip_kernels.hip", tmp_path).items() if "k_strip" in n}
    assert ks
    for name, r in ks.items():
        assert r["ScratchSize"] == 0, (name, r)
