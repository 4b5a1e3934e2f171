"""CPU pins of host/device helpers of the kernels' header
(uni_slam_amd/csrc/orbfe_device.h), built for the host with g++."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

HIP_INCLUDE = "/opt/rocm/include"


def _run(tmp_path, body: str) -> bytes:
    if shutil.which("g++") is None or not os.path.isdir(HIP_INCLUDE):
        pytest.skip("g++ or the HIP headers are absent")
    src = tmp_path / "probe.cpp"
    exe = tmp_path / "probe"
    src.write_text('#include "orbfe_device.h"\n#include <cstdio>\n' + body)
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", HIP_INCLUDE,
                    "-I", os.path.join(ROOT, "uni_slam_amd", "csrc"), "-o", str(exe), str(src)],
                   check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True).stdout


def test_ring_streak_equals_the_shader_loop_for_every_ring(tmp_path):
    """ring_streak(m) == the streak loop of detect_corners.wgsl:67-79 (curr and
    streak start at 1; a set ring bit extends the run, a clear one restarts
    it; no wrap) for all 2^16 ring masks."""
    out = _run(tmp_path, """
int main() {
    for (uint32_t m = 0; m < 65536; ++m) {
        const unsigned char v = static_cast<unsigned char>(orbfe::ring_streak(m));
        fwrite(&v, 1, 1, stdout);
    }
}
""")
    got = np.frombuffer(out, dtype=np.uint8)
    assert got.size == 65536
    m = np.arange(65536, dtype=np.uint32)
    curr = np.ones(65536, np.uint32)
    streak = np.ones(65536, np.uint32)
    for i in range(16):
        bit = (m >> i) & 1
        curr = np.where(bit == 1, curr + 1, 1)
        streak = np.maximum(streak, curr)
    assert np.array_equal(got, streak.astype(np.uint8))
    assert got[0] == 1 and got[0xFFFF] == 17 and got[0b1011] == 3
