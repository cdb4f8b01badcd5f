"""Pins of the oracle's bounce >= 1 estimator (CPU): FP64 closed forms of
plate scenes (plate_closed_forms.py) against the oracle's bake in both
random-stream modes, plus checks of the closed forms themselves.

What the cases pin (reference src/baker.cpp:70-149, 245-294;
src/light.cpp:16-293; src/light.h:72-203; src/material.cpp:113-162):
  point_floor       vertex NEE at depth 0 (radiance x cos x albedo / pi)
  spot_down_plates  continuation weight brdf cos / pdf at depths 0 and 1,
                    vertex NEE at depth 2, the spot cone at a vertex
  spot_up_plates    spot texel NEE (deterministic at 0 bounces), depth-1 and
                    depth-3 vertex NEE through the continuations
  area_near         area-light texel NEE: uniform triangle pick, per-
                    triangle pdf, no emitter cosine, unnormalised projection;
                    depth-0 emission
  area_far          the same light out of reach: its shadow rays end on the
                    emitter itself (tfar = |d| - 0.001 in units of |d|)
Each case is shown to fail when the oracle carries the mistake it targets
(planted on the CPU through orc_set_fault)."""
import math

import numpy as np
import pytest

import plate_closed_forms as pcf
from oracle import oracle as orc

N_TEXELS, SPP = 512, 256
CASES = pcf.case_specs()
PARAMS = [(name, B) for name, c in CASES.items() for B in c["bounces"]]


def expected(name, B):
    sh, vis = CASES[name]["closed"](B)
    sh = np.asarray(sh, np.float64)
    if sh.ndim == 1:
        sh = np.repeat(sh[:, None], 3, 1)
    return sh, vis


def agreement(sh, vis, spp, exp_sh, exp_vis):
    """(worst |mean - expected| / tolerance of the SH, of the visibility):
    the tolerance is 4.5 standard errors of the mean over texels (all texels
    bake the same point, each from its own random numbers), at least 2e-5 of
    the largest coefficient (float rounding of a deterministic term)."""
    m = sh.reshape(-1, 9, 3).astype(np.float64)
    M = m.mean(0)
    se = m.std(0, ddof=1) / math.sqrt(len(m))
    tol = np.maximum(4.5 * se, 2e-5 * np.abs(exp_sh).max())
    p = exp_vis
    vtol = max(4.5 * math.sqrt(p * (1 - p) / (len(m) * spp)), 1e-6)
    return float((np.abs(M - exp_sh) / tol).max()), float(abs(vis.mean() - exp_vis) / vtol)


_oracles = {}


def oracle_of(name):
    if name not in _oracles:
        _oracles[name] = orc.Oracle(CASES[name]["scene"]())
    return _oracles[name]


def bake(name, B, mode, seed=7):
    c = CASES[name]
    pts, w, h = pcf.texel_surface(c["P"], c["N"], N_TEXELS)
    sh, vis, _, _ = oracle_of(name).bake(pts, w, h, SPP, B, mode=mode, seed=seed)
    return sh, vis


@pytest.mark.parametrize("mode", [orc.MODE_REF, orc.MODE_GPU])
@pytest.mark.parametrize("name,B", PARAMS)
def test_oracle_matches_closed_form(name, B, mode):
    exp_sh, exp_vis = expected(name, B)
    sh, vis = bake(name, B, mode)
    a_sh, a_vis = agreement(sh, vis, SPP, exp_sh, exp_vis)
    assert a_sh <= 1.0, (name, B, mode, a_sh)
    assert a_vis <= 1.0, (name, B, mode, a_vis)


# the case each planted mistake must fail (the check fails at > 1)
FAULTS = [(orc.FAULT_CONT_NO_COSINE, "spot_down_plates", 2),
          (orc.FAULT_CONT_NO_COSINE, "spot_up_plates", 3),
          (orc.FAULT_CONT_NO_COSINE, "spot_up_plates", 1),
          (orc.FAULT_NEE_NO_COSINE, "point_floor", 0),
          (orc.FAULT_SPOT_NO_FALLOFF, "spot_down_plates", 1),
          (orc.FAULT_SPOT_NO_FALLOFF, "spot_up_plates", 0),
          (orc.FAULT_AREA_PDF_TOTAL, "area_near", 1)]


@pytest.mark.parametrize("fault,name,B", FAULTS)
def test_closed_forms_catch_planted_mistakes(fault, name, B):
    exp_sh, exp_vis = expected(name, B)
    orc.set_fault(fault)
    try:
        sh, vis = bake(name, B, orc.MODE_REF)
    finally:
        orc.set_fault(0)
    a_sh, _ = agreement(sh, vis, SPP, exp_sh, exp_vis)
    assert a_sh > 1.5, (fault, name, B, a_sh)


@pytest.mark.parametrize("name,B", PARAMS)
def test_closed_forms_converged(name, B, monkeypatch):
    """Doubling every node count moves no coefficient by 1e-7 of the largest."""
    a, va = expected(name, B)
    orig_fields, orig_floor = pcf.Plates.fields, pcf.floor_field_about_light
    monkeypatch.setattr(pcf.Plates, "fields",
                        lambda self, l, b, n=320, n_smooth=640: orig_fields(self, l, b, n,
                                                                             n_smooth))
    monkeypatch.setattr(pcf, "floor_field_about_light",
                        lambda O, foot, F, R, n_phi=1024: orig_floor(O, foot, F, R, n_phi))
    orig_rule = pcf.triangle_rule
    monkeypatch.setattr(pcf, "triangle_rule", lambda a_, b_, c_, n=96: orig_rule(a_, b_, c_, n))
    b, vb = expected(name, B)
    assert np.abs(a - b).max() <= 1e-7 * np.abs(a).max()
    assert va == vb


def test_closed_form_building_blocks():
    # the kernel of a continuation between parallel plates integrates to pi
    # (a uniform plate of radiance 1 returns albedo x 1)
    r, w = pcf.smooth_nodes(1.3, 400)
    one = pcf.Field(r, w, np.ones_like(r))
    for t in (0.0, 0.7, 5.0):
        np.testing.assert_allclose(pcf.convolve([one], np.array([t]), 1.3), np.pi, rtol=1e-9)
    # the floor integral of Y over the square's solid angle, against the
    # closed-form solid angle (Y_0 is constant) and the symmetry of Y_1..
    R, h = 3.0, 1.1
    rr, ww = pcf.piece_nodes([0.0, 1.0, 2.0, 3.0, 3.0 * math.sqrt(2)], 200)
    F = pcf.Field(rr, ww, np.ones_like(rr))
    sh = pcf.floor_field_about_light(np.array([0.0, h, 0.0]), np.zeros(3), F, R, n_phi=2048)
    omega = pcf.square_solid_angle(h, R)
    np.testing.assert_allclose(sh[0], pcf.K[0] * omega, rtol=2e-4)  # square edges cut nodes
    assert abs(sh[3]) < 1e-9 and abs(sh[2]) < 1e-9
    # triangle solid angle against the area quadrature of its solid angle
    O = np.array([0.1, 0.0, -0.2])
    tri = ((-0.4, 0.6, -0.3), (0.1, 0.6, 0.4), (0.3, 0.6, -0.35))
    pts, wts = pcf.triangle_rule(*tri)
    v = pts - O
    dist = np.linalg.norm(v, axis=1)
    np.testing.assert_allclose((wts * v[:, 1] / dist ** 3).sum(),
                               pcf.triangle_solid_angle(O, *tri), rtol=1e-10)
    # sRGB decode as the bake does it (the oracle's table)
    for b in (0, 10, 128, pcf.FLOOR_BYTE, 255):
        assert pcf.srgb_to_linear(b) == orc.srgb_to_linear(b)


@pytest.mark.parametrize("n", [2, 4])
@pytest.mark.parametrize("name,B", [("point_floor", 1), ("spot_down_plates", 2),
                                    ("spot_up_plates", 3), ("area_near", 1)])
def test_oracle_several_light_samples_match_closed_form(name, B, n):
    """num_light_samples = n (light.cpp:164-188, 266-290): n picks per NEE
    and their mean, in both streams; the expectation does not depend on n."""
    exp_sh, exp_vis = expected(name, B)
    c = CASES[name]
    pts, w, h = pcf.texel_surface(c["P"], c["N"], N_TEXELS)
    for mode in (orc.MODE_REF, orc.MODE_GPU):
        sh, vis, _, _ = oracle_of(name).bake(pts, w, h, SPP, B, mode=mode, seed=7,
                                             num_light_samples=n)
        a_sh, a_vis = agreement(sh, vis, SPP, exp_sh, exp_vis)
        assert a_sh <= 1.0 and a_vis <= 1.0, (name, B, n, mode, a_sh)
