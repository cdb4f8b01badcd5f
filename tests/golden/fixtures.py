"""Loading the committed golden fixtures (written by make_golden.py)."""
import os

import numpy as np

from shb import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def scene(z):
    """The fixture's scene as a shb.scenes.Scene."""
    s = scenes.Scene()
    s.positions = z["scene_positions"]
    s.normals = z["scene_normals"]
    s.texture_uvs = z["scene_texture_uvs"]
    s.lightmap_uvs = z["scene_lightmap_uvs"]
    s.tangents = z["scene_tangents"]
    s.indices = z["scene_indices"]
    s.geometries = [dict(vertex_offset=int(g[0]), vertex_count=int(g[1]), index_offset=int(g[2]),
                         index_count=int(g[3]), material_id=int(g[4]), flags=int(g[5]),
                         transform=tuple(float(x) for x in t))
                    for g, t in zip(z["scene_geometries"], z["scene_transforms"])]
    s.materials = [dict(width=int(m[0]), height=int(m[1]), channels=int(m[2]), offset=int(m[3]),
                        emission=float(e))
                   for m, e in zip(z["scene_materials"], z["scene_emission"])]
    s.texels = z["scene_texels"]
    s.lights = [dict(type=int(l["type"]), position=tuple(l["position"]),
                     direction=tuple(l["direction"]), color=tuple(l["color"]),
                     intensity=float(l["intensity"]), cos_inner_cone=float(l["cos_inner_cone"]),
                     cos_outer_cone=float(l["cos_outer_cone"]),
                     geometry_index=int(l["geometry_index"]),
                     material_index=int(l["material_index"]))
                for l in z["scene_lights"]]
    return s


def surface_digest(pts):
    import hashlib
    v = pts[pts["material_id"] >= 0]
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(v).tobytes()).digest(), np.uint8)


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()
