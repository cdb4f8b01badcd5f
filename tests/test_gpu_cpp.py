"""The reference's unit tests of the bake path (baker_test.cpp,
occlusion_test.cpp, rasterizer_test.cpp, dilation_test.cpp, the lightmap
cases of saver_test.cpp), restated in C++ against include/shb_baker.hpp and
run on the GPU (tests/cpp/). Without a GPU the host-only (saver) cases run."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "shb_baker_test")


@pytest.mark.gpu
def test_reference_unit_tests_on_gpu():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


def test_cpp_mirror_builds_and_detects_the_device():
    # without a GPU the host-only cases run and a pass exits 77; with one
    # every case runs
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode in (0, 77), r.stdout + r.stderr
    assert "SaverTest.SavePackedLuminance" in r.stdout and " 0 failed" in r.stdout
