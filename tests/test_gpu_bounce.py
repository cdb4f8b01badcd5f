"""GPU pins of the bounce >= 1 estimator: the HIP bake of the plate and
area-light scenes against their FP64 closed forms (plate_closed_forms.py),
within 4.5 standard errors of the mean over 4096 texels that bake the same
point, and against the oracle's same-stream restatement texel by texel.
The cases and what each pins are listed in test_oracle_bounce.py (Trace,
src/baker.cpp:70-149; EvaluateLightSamples / AccumulateIncomingLightSamples,
src/light.cpp:82-293; light.h:72-203; SampleAreaLight, light.cpp:16-78)."""
import numpy as np
import pytest

import plate_closed_forms as pcf
from oracle import oracle as orc
from test_oracle_bounce import CASES, PARAMS, agreement, expected

pytestmark = pytest.mark.gpu

N_GPU, SPP = 4096, 256


@pytest.mark.parametrize("name,B", PARAMS)
def test_gpu_matches_closed_form(baker, name, B):
    c = CASES[name]
    print(f"closed form {name} bounces {B}", flush=True)
    baker.upload_scene(c["scene"]())
    pts, w, h = pcf.texel_surface(c["P"], c["N"], N_GPU)
    baker.set_surface(pts, w, h, 1)
    sh, vis, st = baker.bake(SPP, B, seed=2024)
    assert st["valid_texels"] == N_GPU
    exp_sh, exp_vis = expected(name, B)
    a_sh, a_vis = agreement(sh, vis, SPP, exp_sh, exp_vis)
    m = sh.reshape(-1, 9, 3).astype(np.float64).mean(0)
    print(f"  worst |GPU - closed form| / tolerance: SH {a_sh:.3f}, visibility {a_vis:.3f}; "
          f"max relative difference {np.abs(m - exp_sh).max() / np.abs(exp_sh).max():.2e}")
    assert a_sh <= 1.0 and a_vis <= 1.0


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_same_stream_as_oracle_on_pinned_scenes(baker, name):
    c = CASES[name]
    B = max(c["bounces"])
    s = c["scene"]()
    baker.upload_scene(s)
    pts, w, h = pcf.texel_surface(c["P"], c["N"], 256)
    baker.set_surface(pts, w, h, 1)
    g_sh, g_vis, _ = baker.bake(16, B, seed=99)
    o_sh, o_vis, _, _ = orc.Oracle(s).bake(pts, w, h, 16, B, mode=orc.MODE_GPU, seed=99)
    scale = np.abs(o_sh).max(axis=1, keepdims=True) + 1e-6
    per_texel = (np.abs(g_sh - o_sh) / scale).max(axis=1)
    assert (per_texel < 1e-3).mean() > 0.97, np.sort(per_texel)[-10:]
    assert np.abs(g_vis - o_vis).max() <= 2.0 / 16


@pytest.mark.parametrize("name,B", [("point_floor", 1), ("spot_down_plates", 2),
                                    ("spot_up_plates", 1), ("area_near", 1)])
def test_gpu_several_light_samples_match_closed_form(baker, name, B):
    """num_light_samples = 4 (light.cpp:164-188, 266-290): four picks per
    NEE, their mean; the expectation is the same as with one."""
    c = CASES[name]
    print(f"closed form {name} bounces {B}, 4 light samples", flush=True)
    baker.upload_scene(c["scene"]())
    pts, w, h = pcf.texel_surface(c["P"], c["N"], N_GPU)
    baker.set_surface(pts, w, h, 1)
    sh, vis, _ = baker.bake(SPP, B, seed=2025, num_light_samples=4)
    exp_sh, exp_vis = expected(name, B)
    a_sh, a_vis = agreement(sh, vis, SPP, exp_sh, exp_vis)
    print(f"  worst |GPU - closed form| / tolerance: SH {a_sh:.3f}, visibility {a_vis:.3f}")
    assert a_sh <= 1.0 and a_vis <= 1.0
