"""bench.py measures one GPU: asked for more it exits with a message before
touching a device. Its VALU issue probe loads and reports three rates."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)


def test_valu_probe_library_exports_its_entry_point():
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "sh-baker_amd", "shb", "libshb_probe.so"))
    assert hasattr(lib, "shb_probe_valu") and hasattr(lib, "shb_probe_cus")


@pytest.mark.gpu
def test_valu_probe_reports_three_issue_rates():
    import json
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--valu-probe"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    line = json.loads(p.stdout.strip().splitlines()[-1])
    kinds = [r["kind"] for r in line["valu_probe"]]
    assert kinds == ["fp32_fma", "int_add_rotate_xor", "mixed_4fma_6int"]
    for r in line["valu_probe"]:
        # at most one wave64 VALU instruction per SIMD per clock
        assert 0.05 < r["valu_per_simd_per_clk"] <= 1.0, r
        assert 500.0 < r["shader_mhz"] < 3000.0, r
