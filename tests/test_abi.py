"""The C-ABI library loads, exports every entry point include/shb.h
declares, and its ctypes mirrors have the C layouts (checked against gcc's
sizeof/offsetof of the header). No GPU compute here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import shb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shb.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(shb_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = shb.get_lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(shb.EXPORTED)


def test_struct_layouts_match_the_header():
    structs = {"shb_geometry": shb.ShbGeometry, "shb_material": shb.ShbMaterial,
               "shb_light": shb.ShbLight, "shb_scene_desc": shb.ShbSceneDesc,
               "shb_raster_cfg": shb.ShbRasterCfg, "shb_bake_cfg": shb.ShbBakeCfg,
               "shb_bake_stats": shb.ShbBakeStats, "shb_tuning": shb.ShbTuning}
    prog = "#include <stdio.h>\n#include <stddef.h>\n#include \"shb.h\"\nint main(){\n"
    for n in structs:
        prog += f'printf("{n} %zu\\n", sizeof({n}));\n'
    prog += 'printf("shb_surface_point %zu\\n", sizeof(shb_surface_point));\n'
    prog += 'printf("shb_ray %zu\\n", sizeof(shb_ray));\n'
    prog += 'printf("shb_hit %zu\\n", sizeof(shb_hit));\n'
    prog += 'printf("off_transform %zu\\n", offsetof(shb_geometry, transform));\n'
    prog += 'printf("off_lights %zu\\n", offsetof(shb_scene_desc, lights));\n'
    prog += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = dict(l.split() for l in subprocess.run([exe], capture_output=True, text=True,
                                                    check=True).stdout.splitlines())
    for n, cls in structs.items():
        assert int(out[n]) == C.sizeof(cls), n
    assert int(out["shb_surface_point"]) == shb.SURFACE_DTYPE.itemsize
    assert int(out["shb_ray"]) == shb.RAY_DTYPE.itemsize
    assert int(out["shb_hit"]) == shb.HIT_DTYPE.itemsize
    assert int(out["off_transform"]) == shb.ShbGeometry.transform.offset
    assert int(out["off_lights"]) == shb.ShbSceneDesc.lights.offset


def test_error_paths_without_device():
    lib = shb.get_lib()
    assert lib.shb_create(None, 0) == -1
    assert lib.shb_upload_scene(None, None) == -1
    assert lib.shb_bake(None, None, None, None, None) == -1
    assert lib.shb_last_error(None) == b"ctx is NULL"
    assert lib.shb_version().startswith(b"shb-hip")


def test_create_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = shb.get_lib().shb_create(C.byref(h), 0)
    assert rc == -3 and not h.value
