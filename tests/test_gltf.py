"""glTF loader (shb_load_gltf, replacing LoadScene, src/loader.cpp:113-844)
on the reference's own fixtures (copied under tests/golden/gltf) and on
small files written here.

Pinned by: the loader_test.cpp KATs (LoadCube, MissingFile,
LoadBoxFallbackColor); an independent reading of the same files with json +
numpy (accessors, degenerate-triangle filter, node transforms as T R S
products, light placement, sun direction); images against Pillow (PNG bit
exact; baseline JPEG within 3 levels of libjpeg, whose IDCT and chroma
upsampling round differently); Radiance RGBE against its defining formula;
generated tangents against the analytic tangent of planar UV mappings. Host
code only: runs on the CPU."""
import base64
import json
import math
import os
import struct

import numpy as np
import pytest

import shb

GL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gltf")
FIXTURES = ["box/scene.gltf", "plane/scene.gltf", "cube/Cube.gltf", "DirectionalLight/scene.gltf",
            "LightsPunctualLamp/LightsPunctualLamp.gltf"]


# ------------------------------------------------------- independent reader
def np_accessor(g, bufs, ai):
    a = g["accessors"][ai]
    bv = g["bufferViews"][a["bufferView"]]
    dt = {5126: "<f4", 5125: "<u4", 5123: "<u2", 5121: "u1"}[a["componentType"]]
    nc = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}[a["type"]]
    esz = np.dtype(dt).itemsize * nc
    stride = bv.get("byteStride", esz)
    base = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
    raw = bufs[bv["buffer"]]
    out = np.empty((a["count"], nc), dt)
    for i in range(a["count"]):
        out[i] = np.frombuffer(raw, dt, nc, base + i * stride)
    return out


def quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def node_mat(n):
    M = np.eye(4)
    if "matrix" in n:
        return np.array(n["matrix"], float).reshape(4, 4).T
    if "translation" in n:
        M[:3, 3] = n["translation"]
    R = quat_mat(n["rotation"]) if "rotation" in n else np.eye(3)
    S = np.diag(n.get("scale", [1, 1, 1]))
    M[:3, :3] = R @ S
    return M


def np_scene(path):
    g = json.load(open(path))
    d = os.path.dirname(path)
    bufs = [open(os.path.join(d, b["uri"]), "rb").read() for b in g["buffers"]]
    geos, lights = [], []
    plights = g.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])

    def walk(ni, parent):
        n = g["nodes"][ni]
        M = parent @ node_mat(n)
        if "mesh" in n:
            for p in g["meshes"][n["mesh"]]["primitives"]:
                at = p["attributes"]
                idx = np_accessor(g, bufs, p["indices"]).ravel().astype(np.uint32)
                t = idx.reshape(-1, 3)
                keep = (t[:, 0] != t[:, 1]) & (t[:, 0] != t[:, 2]) & (t[:, 1] != t[:, 2])
                geos.append(dict(P=np_accessor(g, bufs, at["POSITION"]),
                                 N=np_accessor(g, bufs, at["NORMAL"]),
                                 UV=np_accessor(g, bufs, at["TEXCOORD_0"]),
                                 T=np_accessor(g, bufs, at["TANGENT"]) if "TANGENT" in at else None,
                                 I=t[keep].ravel(), M=M, mat=p["material"]))
        li = n.get("extensions", {}).get("KHR_lights_punctual", {}).get("light")
        if li is not None:
            L = plights[li]
            R = M[:3, :3]
            U, _, Vt = np.linalg.svd(R)
            Rot = U @ Vt
            dvec = Rot @ [0, 0, -1.0]
            lights.append(dict(type=L["type"], position=M[:3, 3], direction=dvec / np.linalg.norm(dvec),
                               intensity=L.get("intensity", 1.0), color=L.get("color", [1, 1, 1])))
        for c in n.get("children", []):
            walk(c, M)

    for ni in g["scenes"][g.get("scene", 0)]["nodes"]:
        walk(ni, np.eye(4))
    return g, geos, lights


@pytest.mark.parametrize("fx", FIXTURES)
def test_fixture_against_independent_reading(fx):
    path = os.path.join(GL, fx)
    s, env = shb.load_gltf(path)
    g, geos, lights = np_scene(path)
    assert len(s.geometries) == len(geos)
    for sg, ng in zip(s.geometries, geos):
        vo, vc, io, ic = sg["vertex_offset"], sg["vertex_count"], sg["index_offset"], sg["index_count"]
        assert vc == len(ng["P"]) and sg["material_id"] == ng["mat"]
        np.testing.assert_array_equal(s.positions[vo:vo + vc], ng["P"])
        np.testing.assert_array_equal(s.normals[vo:vo + vc], ng["N"])
        np.testing.assert_array_equal(s.texture_uvs[vo:vo + vc], ng["UV"])
        if ng["T"] is not None:
            np.testing.assert_array_equal(s.tangents[vo:vo + vc], ng["T"])
        np.testing.assert_array_equal(s.indices[io:io + ic], ng["I"])
        M = np.array(sg["transform"], float).reshape(3, 4)
        np.testing.assert_allclose(M, ng["M"][:3], atol=2e-6)
        assert sg["flags"] & shb.GEOM_HAS_TANGENTS and not sg["flags"] & shb.GEOM_HAS_LIGHTMAP_UVS
    punctual = [l for l in s.lights if l["type"] != 3]
    assert len(punctual) == len(lights)
    kinds = {"point": 0, "directional": 1, "spot": 2}
    for a, b in zip(punctual, lights):
        assert a["type"] == kinds[b["type"]]
        np.testing.assert_allclose(a["position"], b["position"], atol=1e-5)
        np.testing.assert_allclose(a["direction"], b["direction"], atol=1e-5)
        assert a["intensity"] == pytest.approx(b["intensity"])
        np.testing.assert_allclose(a["color"], b["color"])
    # environment: the brightest directional light's reverse direction
    dirs = [l for l in lights if l["type"] == "directional"]
    if dirs:
        best = max(dirs, key=lambda l: l["intensity"])
        assert env["type"] == shb.ENV_PREETHAM
        np.testing.assert_allclose(env["sun_direction"], -best["direction"], atol=1e-5)
    else:
        assert env["type"] == shb.ENV_NONE


def test_load_cube_kat():  # loader_test.cpp:9-50
    s, _ = shb.load_gltf(os.path.join(GL, "cube", "Cube.gltf"))
    assert len(s.geometries) == 1
    g = s.geometries[0]
    assert g["index_count"] == 36 and g["vertex_count"] > 0
    assert len(s.materials) == 1 and s.materials[0]["name"] == "Cube"
    m = s.materials[0]
    assert m["width"] > 0 and m["height"] > 0 and m["channels"] == 4
    assert os.path.isabs(m["albedo_path"]) and m["albedo_path"].endswith("Cube_BaseColor.png")


def test_missing_file():  # loader_test.cpp:52-55
    with pytest.raises(shb.ShbError):
        shb.load_gltf("non_existent.gltf")


def test_box_fallback_color():  # loader_test.cpp:57-82
    s, _ = shb.load_gltf(os.path.join(GL, "box", "scene.gltf"))
    assert len(s.materials) == 1 and s.materials[0]["name"] == "Red"
    m = s.materials[0]
    assert (m["width"], m["height"], m["channels"]) == (1, 1, 4)
    px = s.texels[m["offset"]:m["offset"] + 4]
    assert abs(int(px[0]) - 231) <= 1 and px[1] == 0 and px[2] == 0 and px[3] == 255


def test_images_against_pillow():
    from PIL import Image
    cases = [("cube/Cube.gltf", 0, "cube/Cube_BaseColor.png", 0),
             ("LightsPunctualLamp/LightsPunctualLamp.gltf", 1,
              "LightsPunctualLamp/material1_basecolor.png", 0),
             ("LightsPunctualLamp/LightsPunctualLamp.gltf", 0,
              "LightsPunctualLamp/material0_basecolor.jpeg", 3)]
    for fx, mi, img, tol in cases:
        s, _ = shb.load_gltf(os.path.join(GL, fx))
        m = s.materials[mi]
        ours = s.texels[m["offset"]:m["offset"] + m["width"] * m["height"] * 4]
        ours = ours.reshape(m["height"], m["width"], 4).astype(int)
        ref = np.asarray(Image.open(os.path.join(GL, img)).convert("RGBA")).astype(int)
        d = np.abs(ours - ref)
        assert d.max() <= tol, (img, d.max())
        assert d.mean() < 0.05, (img, d.mean())


# ------------------------------------------------------ files written here
def quad_gltf(tmp_path, uv, with_tangent=False, uv1=False, emissive=0.0, extra=None,
              glb=False, degenerate=False, light=None):
    P = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], np.float32)
    N = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    UV = np.asarray(uv, np.float32)
    I = np.array([0, 1, 2, 0, 2, 3] + ([1, 1, 2] if degenerate else []), np.uint16)
    parts = [P, N, UV]
    if uv1:
        parts.append(UV * 0.5 + 0.25)
    if with_tangent:
        parts.append(np.tile(np.array([[1, 0, 0, 1]], np.float32), (4, 1)))
    blob = b"".join(p.tobytes() for p in parts) + I.tobytes()
    blob += b"\0" * (-len(blob) % 4)
    views, accs, off = [], [], 0
    types = ["VEC3", "VEC3", "VEC2"] + (["VEC2"] if uv1 else []) + (["VEC4"] if with_tangent else [])
    for p, t in zip(parts, types):
        views.append({"buffer": 0, "byteOffset": off, "byteLength": p.nbytes})
        accs.append({"bufferView": len(views) - 1, "componentType": 5126, "count": 4, "type": t})
        off += p.nbytes
    views.append({"buffer": 0, "byteOffset": off, "byteLength": I.nbytes})
    accs.append({"bufferView": len(views) - 1, "componentType": 5123, "count": len(I),
                 "type": "SCALAR"})
    attrs = {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}
    k = 3
    if uv1:
        attrs["TEXCOORD_1"] = k
        k += 1
    if with_tangent:
        attrs["TANGENT"] = k
    mat = {"name": "m", "pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 1.0, 1.0]}}
    if emissive:
        mat["extensions"] = {"KHR_materials_emissive_strength": {"emissiveStrength": emissive}}
    nodes = [{"mesh": 0, "translation": [1, 2, 3], "scale": [2, 2, 2]}]
    g = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}], "nodes": nodes,
         "meshes": [{"primitives": [{"attributes": attrs, "indices": len(accs) - 1,
                                     "material": 0}]}],
         "materials": [mat], "accessors": accs, "bufferViews": views,
         "buffers": [{"byteLength": len(blob)}]}
    if light:
        g["extensions"] = {"KHR_lights_punctual": {"lights": [light]}}
        nodes.append({"extensions": {"KHR_lights_punctual": {"light": 0}},
                      "rotation": [0, 0.7071067811865476, 0, 0.7071067811865476],
                      "translation": [5, 6, 7]})
        g["scenes"][0]["nodes"].append(1)
    if extra:
        g["extras"] = extra
    if glb:
        js = json.dumps(g).encode()
        js += b" " * (-len(js) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        path = tmp_path / "q.glb"
        path.write_bytes(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
    else:
        g["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
        path = tmp_path / "q.gltf"
        path.write_text(json.dumps(g))
    return str(path)


UV_PLAIN = [[0, 0], [1, 0], [1, 1], [0, 1]]
UV_MIRROR = [[1, 0], [0, 0], [0, 1], [1, 1]]


@pytest.mark.parametrize("glb", [False, True])
def test_data_uri_glb_transform_and_fallback_colour(tmp_path, glb):
    s, env = shb.load_gltf(quad_gltf(tmp_path, UV_PLAIN, glb=glb))
    g = s.geometries[0]
    M = np.array(g["transform"]).reshape(3, 4)
    np.testing.assert_allclose(M, [[2, 0, 0, 1], [0, 2, 0, 2], [0, 0, 2, 3]])
    assert g["index_count"] == 6
    # linear (0.5, 0.25, 1.0) -> sRGB8 by src/colorspace.h:17-22
    srgb = [round((1.055 * v ** (1 / 2.4) - 0.055) * 255) for v in (0.5, 0.25, 1.0)]
    assert list(s.texels[:4]) == srgb + [255]
    assert env["type"] == shb.ENV_NONE


@pytest.mark.parametrize("uv,tx,w", [(UV_PLAIN, 1.0, 1.0), (UV_MIRROR, -1.0, -1.0)])
def test_generated_tangents(tmp_path, uv, tx, w):
    # planar quad, u along +x (or -x when mirrored), v along +y: the tangent
    # is dP/du = (tx, 0, 0), unit and orthogonal to n = +z; w is the sign of
    # (n x t) . dP/dv, so the bitangent is w (n x t)
    s, _ = shb.load_gltf(quad_gltf(tmp_path, uv))
    T = s.tangents
    np.testing.assert_allclose(T[:, :3], np.tile([tx, 0, 0], (4, 1)), atol=1e-6)
    np.testing.assert_array_equal(T[:, 3], np.full(4, w))
    # a given TANGENT attribute is kept as is
    s2, _ = shb.load_gltf(quad_gltf(tmp_path, uv, with_tangent=True))
    np.testing.assert_array_equal(s2.tangents, np.tile([1, 0, 0, 1], (4, 1)))


def test_tangent_handedness_with_flipped_v(tmp_path):
    # v running against the bitangent direction flips w
    s, _ = shb.load_gltf(quad_gltf(tmp_path, [[0, 1], [1, 1], [1, 0], [0, 0]]))
    np.testing.assert_allclose(s.tangents[:, :3], np.tile([1, 0, 0], (4, 1)), atol=1e-6)
    np.testing.assert_array_equal(s.tangents[:, 3], -np.ones(4))


def test_lightmap_uvs_emission_degenerate_and_lights(tmp_path):
    light = {"type": "spot", "intensity": 7.5, "color": [1, 0.5, 0.25],
             "spot": {"innerConeAngle": 0.2, "outerConeAngle": 0.6}}
    s, env = shb.load_gltf(quad_gltf(tmp_path, UV_PLAIN, uv1=True, emissive=4.0,
                                     degenerate=True, light=light))
    g = s.geometries[0]
    assert g["flags"] & shb.GEOM_HAS_LIGHTMAP_UVS
    np.testing.assert_allclose(s.lightmap_uvs, np.asarray(UV_PLAIN) * 0.5 + 0.25)
    assert g["index_count"] == 6  # the (1, 1, 2) triangle is dropped
    area = [l for l in s.lights if l["type"] == 3]
    assert len(area) == 1 and area[0]["geometry_index"] == 0 and area[0]["material_index"] == 0
    assert area[0]["intensity"] == 4.0 and s.materials[0]["emission"] == 4.0
    spot = [l for l in s.lights if l["type"] == 2][0]
    np.testing.assert_allclose(spot["position"], [5, 6, 7])
    # +90 degrees about y turns -z into -x
    np.testing.assert_allclose(spot["direction"], [-1, 0, 0], atol=1e-6)
    assert spot["cos_inner_cone"] == pytest.approx(math.cos(0.2), rel=1e-6)
    assert spot["cos_outer_cone"] == pytest.approx(math.cos(0.6), rel=1e-6)
    np.testing.assert_allclose(spot["color"], [1, 0.5, 0.25])
    assert env["type"] == shb.ENV_NONE  # a spot light is not a sun


def test_missing_normal_fails(tmp_path):
    p = quad_gltf(tmp_path, UV_PLAIN)
    g = json.load(open(p))
    del g["meshes"][0]["primitives"][0]["attributes"]["NORMAL"]
    open(p, "w").write(json.dumps(g))
    with pytest.raises(shb.ShbError) as e:
        shb.load_gltf(p)
    assert "NORMAL" in str(e.value)


def write_hdr(path, rgbe, w, h):
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(f"-Y {h} +X {w}\n".encode())
        f.write(np.ascontiguousarray(rgbe, np.uint8).tobytes())


def test_hdr_and_ldr_skybox(tmp_path):
    rng = np.random.default_rng(2)
    w, h = 5, 3  # below 8 columns: flat (not run-length) scanlines
    rgbe = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    rgbe[0, 0, 3] = 0
    write_hdr(tmp_path / "sky.hdr", rgbe, w, h)
    _, env = shb.load_gltf(quad_gltf(tmp_path, UV_PLAIN, extra={"skybox": "sky.hdr"}))
    assert env["type"] == shb.ENV_TEXTURE and env["path"].endswith("sky.hdr")
    e = rgbe[..., 3].astype(np.int64)
    exp = rgbe[..., :3].astype(np.float64) * np.ldexp(1.0, e - 136)[..., None]
    exp[e == 0] = 0
    np.testing.assert_array_equal(env["pixels"], exp.astype(np.float32))
    from PIL import Image
    ldr = rng.integers(0, 256, (4, 6, 3)).astype(np.uint8)
    Image.fromarray(ldr).save(tmp_path / "sky.png")
    _, env = shb.load_gltf(quad_gltf(tmp_path, UV_PLAIN, extra={"skybox": "sky.png"}))
    np.testing.assert_array_equal(env["pixels"], (ldr / np.float32(255.0)).astype(np.float32))


def test_hdr_run_length_scanlines(tmp_path):
    # adaptive RLE: 2 2 w_hi w_lo, then per channel runs (>128: repeat)
    w, h = 16, 2
    rows = []
    exp = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        chans = [np.full(w, 10 + y, np.uint8), np.arange(w, dtype=np.uint8),
                 np.full(w, 200, np.uint8), np.full(w, 129, np.uint8)]
        row = bytes([2, 2, 0, w])
        for c, v in enumerate(chans):
            if c == 1:
                row += bytes([w]) + v.tobytes()       # literal run
            else:
                row += bytes([128 + w, int(v[0])])  # repeated run
        rows.append(row)
        f = np.ldexp(1.0, 129 - 136)
        exp[y] = np.stack([chans[0] * f, chans[1] * f, chans[2] * f], -1)
    with open(tmp_path / "sky.hdr", "wb") as fh:
        fh.write(b"#?RGBE\n\n" + f"-Y {h} +X {w}\n".encode() + b"".join(rows))
    _, env = shb.load_gltf(quad_gltf(tmp_path, UV_PLAIN, extra={"skybox": "sky.hdr"}))
    np.testing.assert_array_equal(env["pixels"], exp)


@pytest.mark.parametrize("field,where,value", [
    ("count", "accessor", 2 ** 62),            # count * components would wrap
    ("count", "accessor", -4),
    ("count", "accessor", 3.5),
    ("byteStride", "view", 2 ** 61),           # (count - 1) * stride would wrap
    ("byteStride", "view", 4),                 # below the element size (12)
    ("byteOffset", "accessor", 2 ** 63),
    ("byteOffset", "view", -8),
    ("byteLength", "view", 10 ** 300),
    ("byteOffset", "view", 120),               # in range, but the view then ends past the buffer
])
def test_malformed_accessor_sizes_fail_cleanly(tmp_path, field, where, value):
    """Sizes from the file are checked before any arithmetic: a crafted count,
    stride or offset fails the load with a message instead of wrapping a
    product and writing past a vector."""
    p = quad_gltf(tmp_path, UV_PLAIN)
    g = json.load(open(p))
    acc = g["accessors"][0]  # POSITION: VEC3 float, 12-byte elements
    target = acc if where == "accessor" else g["bufferViews"][acc["bufferView"]]
    target[field] = value
    open(p, "w").write(json.dumps(g))
    with pytest.raises(shb.ShbError) as e:
        shb.load_gltf(p)
    assert e.value.code == -1 and str(e.value)


def test_accessor_without_buffer_view_is_bounded(tmp_path):
    p = quad_gltf(tmp_path, UV_PLAIN)
    g = json.load(open(p))
    del g["accessors"][0]["bufferView"]
    g["accessors"][0]["count"] = 2 ** 31
    open(p, "w").write(json.dumps(g))
    with pytest.raises(shb.ShbError) as e:
        shb.load_gltf(p)
    assert "too many elements" in str(e.value)
