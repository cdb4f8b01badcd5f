"""The reference's unit tests of the bake path (baker_test.cpp,
occlusion_test.cpp, rasterizer_test.cpp, dilation_test.cpp), restated in C++
against include/shb_baker.hpp and run on the GPU (tests/cpp/)."""
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
    # on a machine without a GPU the program reports SKIP (77); with one it runs
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode in (0, 77), r.stdout + r.stderr
