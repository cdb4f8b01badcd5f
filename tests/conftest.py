import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sh-baker_amd"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # CPU-side artefacts a fresh checkout lacks: build them here (host
    # compilers only; on a GPU box the prebuilt files travel with the tree).
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    shb_dir = os.path.join(ROOT, "sh-baker_amd", "shb")
    if not all(os.path.exists(os.path.join(shb_dir, n))
               for n in ("libshb_hip.so", "libshb_probe.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "sh-baker_amd"), "-j8"],
                       check=True)
    if not os.path.exists(os.path.join(ROOT, "tests", "cpp", "shb_baker_test")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)


@pytest.fixture(scope="session")
def baker():
    import shb
    b = shb.Baker(0)
    yield b
    b.close()
