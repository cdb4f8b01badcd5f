"""FP64 closed forms of the bake on plate scenes, for the bounce >= 1 pins.

The expectation of the reference estimator (src/baker.cpp:70-149, 245-294;
src/light.cpp:82-293; src/light.h:72-203; src/material.cpp:113-162) is
computed here by deterministic quadrature, independently of the oracle and
of the HIP kernels, on scenes where it reduces to low-dimensional integrals:

* a floor plate (y = 0, normal +y) and optionally a ceiling plate
  (y = Hc, normal -y), both squares of half-size R, Lambertian;
* one point or spot light (spot axis straight down or up), or one area light
  made of coplanar triangles;
* every texel at the same point P with the same normal (shb_set_surface).

Texel estimator per sample: SH += L_nee Y(d) + 2 pi Trace(w) Y(w), w uniform
on the texel's hemisphere, so E[SH_k] = nee_k + int Trace(w) Y_k(w) dw.

Trace on the plates: a vertex at depth d on plate s returns its NEE value
N_s (BRDF rho_s/pi, receiver cosine, light radiance; one light, so the pick
probability and the joint pdf are 1) plus, when d < bounces, the cosine-
sampled continuation, whose expectation is (rho_s/pi) int cos L_{d+1} dw.
Between parallel plates at gap D that integral is a convolution over the
other plate with K_D(s) = D^2 / (s^2 + D^2)^2 (cos at both ends D/r, solid
angle D dA / r^3). The light sits on the plates' common axis of symmetry
(the vertical through it), so every field is radial about that axis, and
the angle of the convolution has a closed form:
  int_0^{2 pi} K_D(|y - x|) dphi = 2 pi D^2 A / ((A - B)(A + B))^{3/2},
  A = r^2 + t^2 + D^2, B = 2 r t  (|y| = r, |x| = t about the axis),
leaving 1-D Gauss-Legendre sums. The reference's offsets are kept: vertex
NEE from x + 0.001 n, continuations from x + 0.001 n (gap Hc - 0.001),
primary rays from P + 0.001 N, texel NEE from P itself.

Left out (each below 1e-8 of the values here): the plates' edges for the
inner convolutions (they are taken as infinite; the fields fall off as
r^-4 and vanish outside the spot cone), the reference's drop of
continuations with cos < 1e-3 pi (horizontal reach > 300 D), and NEE
weights below 1e-6 (only beyond about 250 m from a point light).
"""
from __future__ import annotations

import math

import numpy as np

K = np.array([0.282095, 0.488603, 0.488603, 0.488603, 1.092548, 1.092548, 0.315392, 1.092548,
              0.546274])  # sh_coeffs.cpp:27-31
EPS = 0.001  # every ray offset of the reference (baker.cpp:122, 257-259; light.h)


def sh9(d):
    """Y_k of (..., 3) directions, not normalised (sh_coeffs.cpp:50-78)."""
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    return np.stack([np.full_like(x, K[0]), K[1] * y, K[2] * z, K[3] * x, K[4] * x * y,
                     K[5] * y * z, K[6] * (3 * z * z - 1), K[7] * x * z, K[8] * (x * x - y * y)],
                    -1)


def gauss(a, b, n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def srgb_to_linear(byte):
    """colorspace.h:10-14 in float, the value the bake's albedo lookup gives."""
    x = np.float32(byte) / np.float32(255.0)
    if x <= np.float32(0.04045):
        return float(x / np.float32(12.92))
    return float(np.float32(np.power((x + np.float32(0.055)) / np.float32(1.055),
                                     np.float32(2.4), dtype=np.float32)))


# ----------------------------------------------------------- radial fields
class Field:
    """A radial function about the light's axis, known at quadrature nodes
    r with weights w such that sum(w f r) ~ int f(r) r dr."""

    def __init__(self, r, w, f):
        self.r, self.w, self.f = r, w, f


def ring(r, t, D):
    """int_0^{2 pi} K_D(|y - x|) dphi for |y| = r, |x| = t."""
    A = r * r + t * t + D * D
    B = 2 * r * t
    return 2 * np.pi * D * D * A / ((A - B) * (A + B)) ** 1.5


def convolve(fields, t, D):
    """(K_D * sum of fields)(t) at radii t (any shape)."""
    t = np.asarray(t, np.float64)
    out = np.zeros(t.shape)
    flat = t.reshape(-1)
    acc = out.reshape(-1)
    for F in fields:
        for s in range(0, flat.size, 4096):
            tt = flat[s:s + 4096, None]
            acc[s:s + 4096] += (ring(F.r[None], tt, D) * (F.w * F.f * F.r)[None]).sum(1)
    return out


def smooth_nodes(scale, n):
    """Nodes on [0, inf) for smooth fields: r = scale tan(beta)."""
    b, wb = gauss(0.0, np.pi / 2, n)
    return scale * np.tan(b), wb * scale / np.cos(b) ** 2


def piece_nodes(breaks, n):
    rs, ws = [], []
    for a, b in zip(breaks[:-1], breaks[1:]):
        if b > a:
            r, w = gauss(a, b, n)
            rs.append(r)
            ws.append(w)
    return np.concatenate(rs), np.concatenate(ws)


# ------------------------------------------------------------------ lights
class Light:
    """Point (cos_i = cos_o = None) or spot light at `pos`, straight down
    (axis -1) or up (axis +1), white, intensity I (light.h:72-158)."""

    def __init__(self, pos, I, axis=-1, cos_inner=None, cos_outer=None):
        self.pos = np.asarray(pos, np.float64)
        self.I = float(I)
        self.axis = axis
        self.ci, self.co = cos_inner, cos_outer

    @property
    def spot(self):
        return self.ci is not None

    def falloff(self, cos_l):
        if not self.spot:
            return np.ones_like(cos_l)
        return np.clip((cos_l - self.co) / (self.ci - self.co), 0.0, 1.0)

    def nee_field(self, plate_y, normal_y, rho, n=160):
        """N_s(r): vertex NEE value on a plate at height plate_y with normal
        (0, normal_y, 0) at horizontal distance r from the light's axis
        (light.h, EvaluateLightSamples with one light: radiance x cos x
        rho / pi); None when the light cannot reach that side."""
        h = (self.pos[1] - (plate_y + EPS * normal_y)) * normal_y  # vertical gap along n
        if h <= 0:
            return None
        if self.spot:
            # cos_l = axis . (point - light)/d: the plate must lie on the axis side
            if (normal_y > 0) != (self.axis < 0):
                return None
            tan_i = math.sqrt(1 - self.ci ** 2) / self.ci
            tan_o = math.sqrt(1 - self.co ** 2) / self.co
            r, w = piece_nodes([0.0, h * tan_i, h * tan_o], n)
        else:
            r, w = smooth_nodes(h, 4 * n)
        d2 = r * r + h * h
        d = np.sqrt(d2)
        cos = h / d
        f = (rho / np.pi) * self.I * self.falloff(cos) * cos / d2
        return Field(r, w, f)

    def texel_nee(self, P, N):
        """Texel NEE (light.cpp:191-293 with one light): deterministic
        radiance I falloff / d^2 projected along the unnormalised d = L - P;
        zero when the light is behind the texel or outside the cone. The
        shadow ray's reach past the light is checked by the caller's scene."""
        Lv = self.pos - np.asarray(P, np.float64)
        d = np.linalg.norm(Lv)
        cos_n = np.dot(N, Lv) / d
        if cos_n < 0:
            return np.zeros(9)
        cos_l = -self.axis * Lv[1] / d if self.spot else 1.0
        rad = self.I * float(self.falloff(np.asarray(cos_l))) / (d * d)
        return rad * sh9(Lv)


# ------------------------------------------------------------- the plates
class Plates:
    """Floor (y = 0) and optional ceiling (y = Hc) of half-size R."""

    def __init__(self, rho_floor, rho_ceiling=None, Hc=None, R=200.0):
        self.rho = {0: rho_floor, 1: rho_ceiling}
        self.Hc, self.R = Hc, R

    def plate(self, depth):
        """Depth d hits the floor for even d (primary rays go down)."""
        return depth % 2

    def fields(self, light, bounces, n=160, n_smooth=320):
        """(continuation part of the depth-0 radiance as a Field on smooth
        nodes or None, the floor's NEE Field or None) by the recursion
        L_d = N_s(d) + [d < B] (rho_s(d) / pi) K_D * L_{d+1}."""
        nee = {0: light.nee_field(0.0, 1.0, self.rho[0], n), 1: None}
        if self.Hc is None or bounces == 0:
            return None, nee[0]
        nee[1] = light.nee_field(self.Hc, -1.0, self.rho[1], n)
        D = self.Hc - EPS  # continuations start 0.001 off their plate
        rs, ws = smooth_nodes(D, n_smooth)
        below = [f for f in (nee[self.plate(bounces)],) if f is not None]  # L_B
        for d in range(bounces - 1, -1, -1):
            s = self.plate(d)
            cont = (self.rho[s] / np.pi) * convolve(below, rs, D) if below else 0 * rs
            if d == 0:
                return Field(rs, ws, cont), nee[0]
            below = [f for f in (nee[s],) if f is not None] + [Field(rs, ws, cont)]


# ------------------------------------------------------- texel integrals
def floor_field_about_light(O, foot, F, R, n_phi=512):
    """int over the floor square [-R, R]^2 of F(|x - foot|) Y(x - O) dw,
    seen from O above the floor, with F a Field on nodes about the light's
    axis (a spot cone's kinks at node breaks). Points off the square (where
    F is below 1e-12 of its peak) are dropped."""
    h = O[1]
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    x = np.stack([foot[0] + F.r[None] * np.cos(ph)[:, None], np.zeros((n_phi, F.r.size)),
                  foot[2] + F.r[None] * np.sin(ph)[:, None]], -1)
    on = (np.abs(x[..., 0]) <= R) & (np.abs(x[..., 2]) <= R)
    v = x - O[None, None]
    dist = np.linalg.norm(v, axis=-1)
    dw = h / dist ** 3  # solid angle per unit floor area
    wgt = np.where(on, (F.w * F.r * F.f)[None] * dw, 0.0) * (2 * np.pi / n_phi)
    return (wgt[..., None] * sh9(v / dist[..., None])).sum((0, 1))


def square_solid_angle(h, R):
    """Solid angle of the square [-R, R]^2 seen from height h above its centre."""
    return 4 * math.asin(R * R / (R * R + h * h))


def triangle_solid_angle(O, a, b, c):
    """Van Oosterom & Strackee (IEEE TBME 1983)."""
    a, b, c = (np.asarray(v, np.float64) - O for v in (a, b, c))
    la, lb, lc = (np.linalg.norm(v) for v in (a, b, c))
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2 * math.atan2(num, den)


def triangle_rule(a, b, c, n=48):
    """Points and weights integrating over triangle abc (Duffy map of a
    Gauss-Legendre square; exact area)."""
    u, wu = gauss(0.0, 1.0, n)
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu)
    s, t = U, V * (1 - U)  # (s, t) in the unit triangle, Jacobian (1 - U)
    a, b, c = (np.asarray(v, np.float64) for v in (a, b, c))
    pts = a[None, None] + s[..., None] * (b - a) + t[..., None] * (c - a)
    area2 = np.linalg.norm(np.cross(b - a, c - a))
    return pts.reshape(-1, 3), (W * (1 - U)).reshape(-1) * area2


# ----------------------------------------------------------------- cases
def plate_case(plates, light, P, N, bounces):
    """Expected SH9 (9 coefficients, one channel: light and albedos are
    grey) and environment visibility of a texel at P, normal N (pointing
    down), over the plates."""
    P = np.asarray(P, np.float64)
    N = np.asarray(N, np.float64)
    assert N[1] < 0, "the texel looks down at the floor"
    O = P + EPS * N  # primary ray origin (baker.cpp:257-259)
    foot = np.array([light.pos[0], 0.0, light.pos[2]])
    sh = light.texel_nee(P, N)
    for F in plates.fields(light, bounces):
        if F is not None:
            sh = sh + floor_field_about_light(O, foot, F, plates.R)
    vis = 1.0 - square_solid_angle(O[1], plates.R) / (2 * np.pi)
    return sh, vis


def area_case(tris, emission, P, N, nee_reaches=True):
    """A texel at P (normal N, +y) below coplanar emitting triangles facing
    down, nothing else: E[SH] = texel NEE + depth-0 emission.
    Texel NEE (light.cpp:16-78, 191-293, light.h:166-188): uniform triangle
    pick, uniform point, pdf 1 / (ntri A_tri), radiance e / |d|^2 without
    the emitter's cosine, projected along the unnormalised d = y - P; its
    expectation is int e Y(y - P) / |y - P|^2 dA over the emitter. When
    every point is farther than 1.001 the shadow ray reaches its own
    triangle (tfar = |d| - 0.001 in units of |d|) and the term is zero
    (nee_reaches False). Emission: primary rays that hit the emitter add
    e (baker.cpp:117-119), int e Y(w) dw over its solid angle. The vertex
    NEE at the emitter is zero (its own light lies in the tangent plane,
    0.001 behind the offset shading point) and continuations escape."""
    P = np.asarray(P, np.float64)
    O = P + EPS * np.asarray(N, np.float64)
    e = np.asarray(emission, np.float64)
    nee = np.zeros(9)
    emit = np.zeros(9)
    omega = 0.0
    for a, b, c in tris:
        pts, w = triangle_rule(a, b, c)
        d = pts - P[None]
        d2 = (d * d).sum(1)
        if nee_reaches:
            nee += ((w / d2)[:, None] * sh9(d)).sum(0)
        v = pts - O[None]
        dist = np.linalg.norm(v, axis=1)
        n = np.cross(np.asarray(b) - np.asarray(a), np.asarray(c) - np.asarray(a))
        n = n / np.linalg.norm(n)
        dw = np.abs(v @ n) / dist ** 3
        emit += ((w * dw)[:, None] * sh9(v / dist[:, None])).sum(0)
        omega += triangle_solid_angle(O, a, b, c)
    sh = nee[:, None] * e[None] + emit[:, None] * e[None]  # (9, 3)
    return sh, 1.0 - omega / (2 * np.pi)


# ------------------------------------------------------ the scenes baked
FLOOR_BYTE, CEILING_BYTE = 228, 200  # sRGB8 grey albedos of the plates


def plates_scene(light, Hc=None, R=200.0):
    """shb Scene of the floor (and ceiling) squares and one light dict."""
    from shb import scenes
    b = scenes.SceneBuilder()
    mf = b.add_material((FLOOR_BYTE,) * 3 + (255,))
    b.add_quads([((-R, 0.0, -R), (0.0, 0.0, 2 * R), (2 * R, 0.0, 0.0))], mf)
    if Hc is not None:
        mc = b.add_material((CEILING_BYTE,) * 3 + (255,))
        b.add_quads([((-R, Hc, -R), (2 * R, 0.0, 0.0), (0.0, 0.0, 2 * R))], mc)
    b.add_light(**light)
    return b.build(64, 64, 0)


def area_scene(tris, rgb_bytes, intensity):
    """Coplanar emitting triangles (one geometry, one area light) facing -y."""
    import shb
    from shb import scenes
    b = scenes.SceneBuilder()
    m = b.add_material(tuple(rgb_bytes) + (255,), emission=float(intensity))
    P = np.concatenate([np.asarray(t, np.float64) for t in tris])
    n = np.repeat(np.array([[0.0, -1.0, 0.0]]), len(P), 0)
    tan = np.repeat(np.array([[1.0, 0.0, 0.0, 1.0]]), len(P), 0)
    idx = np.arange(len(P))
    cid, loc = b.new_chart(P[:, [0, 2]])
    g = b.add_geometry(P, n, tan, idx, np.full(len(P), cid), loc, m)
    b.add_light(type=shb.LIGHT_AREA, geometry_index=g, material_index=m, intensity=float(intensity))
    return b.build(64, 64, 0)


def texel_surface(P, N, n_texels, width=64):
    """n_texels copies of one surface point (material 0), as a width-wide grid."""
    import shb
    h = n_texels // width
    pts = np.zeros(width * h, shb.SURFACE_DTYPE)
    pts["position"] = P
    pts["normal"] = N
    pts["tangent"] = (1.0, 0.0, 0.0, 1.0)
    pts["material_id"] = 0
    return pts, width, h


# light dicts and texel placements of the pinned cases
EMITTER_BYTES, EMITTER_INTENSITY = (255, 200, 128), 3.0


def emitter_tris(height):
    """Two coplanar triangles of different areas at y = height, facing -y."""
    a, b, c, d = (-0.4, height, -0.3), (0.3, height, -0.35), (0.1, height, 0.4), \
        (0.5, height, 0.1)
    return [(a, c, b), (b, c, d)]


def case_specs():
    """name -> dict(scene: () -> Scene, closed: bounces -> (sh (9,) grey or
    (9, 3) RGB, visibility), P, N, bounces: the bounce counts pinned)."""
    import shb
    rho_f, rho_c = srgb_to_linear(FLOOR_BYTE), srgb_to_linear(CEILING_BYTE)
    down, up = (0.0, -1.0, 0.0), (0.0, 1.0, 0.0)
    P = (0.0, 1.2, 0.0)
    cases = {}
    # floor lit by a point light; texel above, looking down: the vertex NEE
    # at depth 0 (light.cpp:82-189 with BRDF x cos); continuations escape
    pa = dict(type=shb.LIGHT_POINT, position=(0.7, 1.5, -0.4), intensity=30.0, color=(1, 1, 1))
    la = Light(pa["position"], pa["intensity"])
    cases["point_floor"] = dict(scene=lambda: plates_scene(pa),
                                closed=lambda B: plate_case(Plates(rho_f), la, P, down, B),
                                P=P, N=down, bounces=(0, 1))
    # two plates, spot straight down lights the floor only: B = 1 is the
    # direct term, B = 2 adds floor -> ceiling -> floor (the continuation
    # weight brdf cos / pdf at depths 0 and 1, vertex NEE at depth 2)
    ci, co = math.cos(math.radians(30)), math.cos(math.radians(50))
    pb = dict(type=shb.LIGHT_SPOT, position=(0.7, 0.6, -0.4), direction=down, intensity=8.0,
              color=(1, 1, 1), cos_inner_cone=ci, cos_outer_cone=co)
    lb = Light(pb["position"], pb["intensity"], -1, float(np.float32(ci)), float(np.float32(co)))
    cases["spot_down_plates"] = dict(
        scene=lambda: plates_scene(pb, Hc=1.5),
        closed=lambda B: plate_case(Plates(rho_f, rho_c, 1.5), lb, P, down, B),
        P=P, N=down, bounces=(1, 2))
    # two plates, spot straight up lights the ceiling and the texel: B = 0
    # is the texel NEE alone (deterministic), B = 1 adds floor -> ceiling,
    # B = 3 floor -> ceiling -> floor -> ceiling
    ci2, co2 = math.cos(math.radians(30)), math.cos(math.radians(55))
    pc = dict(type=shb.LIGHT_SPOT, position=(0.25, 0.9, -0.15), direction=up, intensity=6.0,
              color=(1, 1, 1), cos_inner_cone=ci2, cos_outer_cone=co2)
    lc = Light(pc["position"], pc["intensity"], +1, float(np.float32(ci2)),
               float(np.float32(co2)))
    cases["spot_up_plates"] = dict(
        scene=lambda: plates_scene(pc, Hc=1.5),
        closed=lambda B: plate_case(Plates(rho_f, rho_c, 1.5), lc, P, down, B),
        P=P, N=down, bounces=(0, 1, 3))
    # a two-triangle area light within reach (every point closer than 1):
    # texel NEE with the uniform triangle pick and per-triangle pdf, plus
    # the depth-0 emission of primary rays that hit it
    e = [srgb_to_linear(b) * EMITTER_INTENSITY for b in EMITTER_BYTES]
    Pa = (0.0, 0.0, 0.0)
    cases["area_near"] = dict(
        scene=lambda: area_scene(emitter_tris(0.6), EMITTER_BYTES, EMITTER_INTENSITY),
        closed=lambda B: area_case(emitter_tris(0.6), e, Pa, up, True),
        P=Pa, N=up, bounces=(1,))
    # the same emitter out of reach (every point farther than 1.001): the
    # texel NEE shadow ray ends on the emitter itself and the term is zero
    cases["area_far"] = dict(
        scene=lambda: area_scene(emitter_tris(1.6), EMITTER_BYTES, EMITTER_INTENSITY),
        closed=lambda B: area_case(emitter_tris(1.6), e, Pa, up, False),
        P=Pa, N=up, bounces=(1,))
    return cases
