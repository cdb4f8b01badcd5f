"""Python bindings of libshb_hip.so (the C ABI in include/shb.h).

The GPU path is the only path: importing this package loads the in-tree
HIP library and raises if it is missing. Host buffers are numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshb_hip.so")

# ------------------------------------------------------------------ structs
SHB_OK = 0
ERRORS = {
    -1: "SHB_ERR_INVALID_ARG",
    -2: "SHB_ERR_NO_SCENE",
    -3: "SHB_ERR_DEVICE",
    -4: "SHB_ERR_OOM",
    -5: "SHB_ERR_LIMIT",
    -6: "SHB_ERR_NO_SURFACE",
    -7: "SHB_ERR_UNSUPPORTED",
}
LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_SPOT, LIGHT_AREA = 0, 1, 2, 3
GEOM_HAS_NORMALS, GEOM_HAS_TEXTURE_UVS, GEOM_HAS_LIGHTMAP_UVS, GEOM_HAS_TANGENTS = 1, 2, 4, 8
BAKE_COUNT_VISITS = 1
BAKE_USE_GRAPH = 2
BAKE_TRACE_ALL_SHADOW = 4
SAVE_COMBINED, SAVE_SPLIT, SAVE_LUMINANCE_PACKED = 0, 1, 2


class ShbGeometry(C.Structure):
    _fields_ = [("vertex_offset", C.c_uint32), ("vertex_count", C.c_uint32),
                ("index_offset", C.c_uint32), ("index_count", C.c_uint32),
                ("material_id", C.c_int32), ("flags", C.c_uint32),
                ("transform", C.c_float * 12)]


class ShbMaterial(C.Structure):
    _fields_ = [("albedo_width", C.c_uint32), ("albedo_height", C.c_uint32),
                ("albedo_channels", C.c_uint32), ("reserved", C.c_uint32),
                ("albedo_offset", C.c_uint64), ("emission_intensity", C.c_float),
                ("reserved2", C.c_float)]


class ShbLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", C.c_float * 3),
                ("direction", C.c_float * 3), ("color", C.c_float * 3),
                ("intensity", C.c_float), ("cos_inner_cone", C.c_float),
                ("cos_outer_cone", C.c_float), ("geometry_index", C.c_int32),
                ("material_index", C.c_int32)]


_fp = C.POINTER(C.c_float)


class ShbSceneDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("positions", _fp), ("normals", _fp),
                ("texture_uvs", _fp), ("lightmap_uvs", _fp), ("tangents", _fp),
                ("num_indices", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("num_geometries", C.c_uint32), ("geometries", C.POINTER(ShbGeometry)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(ShbMaterial)),
                ("texel_pool_size", C.c_uint64), ("texel_pool", C.POINTER(C.c_uint8)),
                ("num_lights", C.c_uint32), ("lights", C.POINTER(ShbLight))]


class ShbEnvironmentDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("channels", C.c_int32), ("pixels", _fp), ("sun_direction", C.c_float * 3),
                ("sun_intensity", C.c_float), ("path", C.c_char_p)]


ENV_NONE, ENV_TEXTURE, ENV_PREETHAM = 0, 1, 2


class ShbRasterCfg(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("supersample_scale", C.c_int32)]


SURFACE_DTYPE = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("tangent", "<f4", 4),
                          ("material_id", "<i4"), ("reserved", "<i4")])
assert SURFACE_DTYPE.itemsize == 48


class ShbBakeCfg(C.Structure):
    _fields_ = [("samples", C.c_int32), ("bounces", C.c_int32),
                ("num_light_samples", C.c_int32), ("flags", C.c_uint32), ("seed", C.c_uint64)]


class ShbBakeStats(C.Structure):
    _fields_ = [("valid_texels", C.c_uint64), ("paths", C.c_uint64),
                ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("seconds_total", C.c_double), ("seconds_device", C.c_double),
                ("seconds_trace", C.c_double),
                ("node_visits_shadow", C.c_uint64), ("tri_tests_shadow", C.c_uint64),
                ("seconds_trace_shadow", C.c_double),
                ("launches_extension", C.c_uint32), ("launches_shadow", C.c_uint32),
                ("branches", C.c_uint64), ("passes", C.c_uint32), ("reserved", C.c_uint32),
                ("rays_shadow_resolved", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ShbTile(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32)]


class ShbTuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("batch_paths", C.c_int32),
                ("queue_chunk", C.c_int32), ("bvh_width", C.c_int32)]


RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tnear", "<f4"), ("direction", "<f4", 3),
                      ("tfar", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim_id", "<i4"),
                      ("geom_id", "<i4"), ("material_id", "<i4"), ("position", "<f4", 3),
                      ("normal", "<f4", 3), ("uv", "<f4", 2)])
assert HIT_DTYPE.itemsize == 56

EXPORTED = [
    "shb_create", "shb_destroy", "shb_last_error", "shb_version", "shb_upload_scene",
    "shb_rasterize", "shb_set_surface", "shb_validity_mask", "shb_bake", "shb_bake_lightmap", "shb_read_lightmap", "shb_save_lightmap",
    "shb_env_sh_latlong", "shb_env_sh_latlong_noalloc", "shb_env_sh_preetham",
    "shb_load_gltf", "shb_gltf_scene", "shb_gltf_environment", "shb_gltf_material_name",
    "shb_gltf_albedo_path", "shb_gltf_free", "shb_atlas_geometry_scales", "shb_atlas_create",
    "shb_atlas_scene", "shb_atlas_info", "shb_atlas_free",
    "shb_downsample", "shb_downsample_mask", "shb_dilate", "shb_trace_rays", "shb_bvh_export",
    "shb_bake_batched", "shb_bake_tiled", "shb_autotune", "shb_set_tuning", "shb_get_tuning", "shb_set_params",
    "shb_bake_adjoint",
]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C sh-baker_amd` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    vp, ip, u8p = C.c_void_p, C.c_int, C.POINTER(C.c_uint8)
    lib.shb_create.argtypes = [C.POINTER(vp), ip]
    lib.shb_destroy.argtypes = [vp]
    lib.shb_destroy.restype = None
    lib.shb_last_error.argtypes = [vp]
    lib.shb_last_error.restype = C.c_char_p
    lib.shb_version.restype = C.c_char_p
    lib.shb_upload_scene.argtypes = [vp, C.POINTER(ShbSceneDesc)]
    lib.shb_rasterize.argtypes = [vp, C.POINTER(ShbRasterCfg), vp]
    lib.shb_set_surface.argtypes = [vp, C.POINTER(ShbRasterCfg), vp]
    lib.shb_validity_mask.argtypes = [vp, u8p]
    lib.shb_bake.argtypes = [vp, C.POINTER(ShbBakeCfg), _fp, _fp, C.POINTER(ShbBakeStats)]
    lib.shb_bake_lightmap.argtypes = [vp, C.POINTER(ShbBakeCfg), C.c_int32, _fp, _fp, u8p,
                                      C.POINTER(ShbBakeStats)]
    lib.shb_read_lightmap.argtypes = [vp, _fp, _fp, u8p]
    lib.shb_save_lightmap.argtypes = [C.c_char_p, C.c_int32, C.c_int32, _fp, _fp, C.c_int32,
                                      C.c_char_p, C.c_size_t]
    lib.shb_env_sh_latlong.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, _fp, _fp,
                                       C.POINTER(C.c_double)]
    lib.shb_env_sh_latlong_noalloc.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp, vp]
    lib.shb_env_sh_preetham.argtypes = [_fp, _fp]
    lib.shb_load_gltf.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_size_t]
    lib.shb_gltf_scene.restype = C.POINTER(ShbSceneDesc)
    lib.shb_gltf_scene.argtypes = [vp]
    lib.shb_gltf_environment.argtypes = [vp, C.POINTER(ShbEnvironmentDesc)]
    lib.shb_gltf_material_name.restype = C.c_char_p
    lib.shb_gltf_material_name.argtypes = [vp, C.c_uint32]
    lib.shb_gltf_albedo_path.restype = C.c_char_p
    lib.shb_gltf_albedo_path.argtypes = [vp, C.c_uint32]
    lib.shb_gltf_free.argtypes = [vp]
    lib.shb_atlas_geometry_scales.argtypes = [C.POINTER(ShbSceneDesc), C.c_float, _fp]
    lib.shb_atlas_create.argtypes = [C.POINTER(ShbSceneDesc), C.c_uint32, C.c_uint32, C.c_float,
                                     C.POINTER(vp), C.c_char_p, C.c_size_t]
    lib.shb_atlas_scene.restype = C.POINTER(ShbSceneDesc)
    lib.shb_atlas_scene.argtypes = [vp]
    lib.shb_atlas_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_uint32), _fp,
                                   C.POINTER(C.POINTER(C.c_uint32)),
                                   C.POINTER(C.POINTER(C.c_uint32))]
    lib.shb_atlas_free.argtypes = [vp]
    lib.shb_downsample.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _fp, _fp]
    lib.shb_downsample_mask.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, u8p, u8p]
    lib.shb_dilate.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, _fp, u8p, C.c_int32]
    lib.shb_trace_rays.argtypes = [vp, vp, C.c_uint32, C.c_int32, vp,
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.shb_bvh_export.argtypes = [vp, vp, C.POINTER(C.c_uint32), vp, C.POINTER(C.c_uint32)]
    lib.shb_bake_batched.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(ShbBakeCfg),
                                     C.POINTER(_fp), C.POINTER(_fp), C.POINTER(ShbBakeStats)]
    lib.shb_autotune.argtypes = [vp, C.POINTER(ShbBakeCfg), C.POINTER(ShbTuning)]
    lib.shb_bake_tiled.argtypes = [vp, C.POINTER(ShbBakeCfg), C.POINTER(ShbTile), C.c_int32,
                                   _fp, _fp, C.POINTER(ShbBakeStats)]
    lib.shb_set_tuning.argtypes = [vp, C.POINTER(ShbTuning)]
    lib.shb_get_tuning.argtypes = [vp, C.POINTER(ShbTuning)]
    lib.shb_set_params.argtypes = [vp, _fp, _fp]
    lib.shb_bake_adjoint.argtypes = [vp, C.POINTER(ShbBakeCfg), _fp, _fp, _fp,
                                     C.POINTER(ShbBakeStats)]
    return lib


lib = None


def get_lib():
    """Loads libshb_hip.so on first use (raises ImportError if it is missing)."""
    global lib
    if lib is None:
        lib = _load()
    return lib


def version() -> str:
    return get_lib().shb_version().decode()


class ShbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _f(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8)) if a is not None else None


class Baker:
    """One shb_ctx: a scene, a surface, and the bake entry points."""

    def __init__(self, device: int = 0):
        get_lib()
        h = C.c_void_p()
        rc = lib.shb_create(C.byref(h), device)
        if rc != SHB_OK:
            raise ShbError(rc, "shb_create failed (no HIP device?)")
        self.h = h
        self.width = self.height = 0
        self.ss = 1
        self.scene = None

    def close(self):
        if self.h:
            lib.shb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != SHB_OK:
            raise ShbError(rc, lib.shb_last_error(self.h).decode())

    # -- scene / surface
    def upload_scene(self, scene):
        desc, keep = scene.to_desc()
        self._chk(lib.shb_upload_scene(self.h, C.byref(desc)))
        self.scene = scene
        del keep

    def rasterize(self, width, height, ss=1, want=True):
        cfg = ShbRasterCfg(width, height, ss)
        out = np.zeros(width * ss * height * ss, SURFACE_DTYPE) if want else None
        self._chk(lib.shb_rasterize(self.h, C.byref(cfg),
                                    out.ctypes.data_as(C.c_void_p) if want else None))
        self.width, self.height, self.ss = width, height, ss
        return out

    def set_surface(self, points, width, height, ss=1):
        points = np.ascontiguousarray(points, SURFACE_DTYPE)
        assert points.size == width * ss * height * ss
        cfg = ShbRasterCfg(width, height, ss)
        self._chk(lib.shb_set_surface(self.h, C.byref(cfg), points.ctypes.data_as(C.c_void_p)))
        self.width, self.height, self.ss = width, height, ss

    def validity_mask(self):
        n = self.width * self.ss * self.height * self.ss
        m = np.zeros(n, np.uint8)
        self._chk(lib.shb_validity_mask(self.h, _u8(m)))
        return m

    # -- bake
    @staticmethod
    def cfg(samples, bounces, seed=12345, flags=0, num_light_samples=1):
        return ShbBakeCfg(samples, bounces, num_light_samples, flags, seed)

    def bake(self, samples, bounces, seed=12345, flags=0, num_light_samples=1, want=True):
        """The supersampled SH9 bake; with want=False the result stays in
        device memory and None is returned for the arrays."""
        n = self.width * self.ss * self.height * self.ss
        sh = np.zeros((n, 27), np.float32) if want else None
        vis = np.zeros(n, np.float32) if want else None
        st = ShbBakeStats()
        cfg = self.cfg(samples, bounces, seed, flags, num_light_samples)
        self._chk(lib.shb_bake(self.h, C.byref(cfg), _f(sh), _f(vis), C.byref(st)))
        return sh, vis, st.as_dict()

    def bake_tiled(self, samples, bounces, tiles, seed=12345, flags=0, want=True,
                   num_light_samples=1):
        """shb_bake_tiled: the surface baked as tile jobs (x0, y0, w, h in
        supersampled texels) of one batched run; returns (sh, vis, stats)
        like bake (texels outside every tile are 0)."""
        n = self.width * self.ss * self.height * self.ss
        sh = np.zeros((n, 27), np.float32) if want else None
        vis = np.zeros(n, np.float32) if want else None
        arr = (ShbTile * max(1, len(tiles)))(*[ShbTile(*map(int, t)) for t in tiles])
        st = ShbBakeStats()
        cfg = self.cfg(samples, bounces, seed, flags, num_light_samples)
        self._chk(lib.shb_bake_tiled(self.h, C.byref(cfg), arr, len(tiles), _f(sh), _f(vis),
                                     C.byref(st)))
        return sh, vis, st.as_dict()

    def bake_lightmap(self, samples, bounces, dilation, seed=12345, flags=0, to_host=True):
        """Bake + downsample + dilation. With to_host=False the result stays
        in device memory (fetch it with read_lightmap) and None is returned
        for the arrays."""
        n = self.width * self.height
        sh = np.zeros((n, 27), np.float32) if to_host else None
        vis = np.zeros(n, np.float32) if to_host else None
        mask = np.zeros(n, np.uint8) if to_host else None
        st = ShbBakeStats()
        cfg = self.cfg(samples, bounces, seed, flags)
        self._chk(lib.shb_bake_lightmap(self.h, C.byref(cfg), dilation, _f(sh), _f(vis),
                                        _u8(mask), C.byref(st)))
        return sh, vis, mask, st.as_dict()

    def read_lightmap(self):
        n = self.width * self.height
        sh = np.zeros((n, 27), np.float32)
        vis = np.zeros(n, np.float32)
        mask = np.zeros(n, np.uint8)
        self._chk(lib.shb_read_lightmap(self.h, _f(sh), _f(vis), _u8(mask)))
        return sh, vis, mask

    def autotune(self, samples, bounces, seed=12345):
        """Times short bakes of the current scene/surface over launch shapes,
        BVH widths and batch sizes; applies and returns the fastest tuning."""
        t = ShbTuning()
        cfg = self.cfg(samples, bounces, seed)
        self._chk(lib.shb_autotune(self.h, C.byref(cfg), C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    # -- image ops
    def downsample(self, img, width, height, scale):
        img = np.ascontiguousarray(img, np.float32)
        ch = img.size // (width * scale * height * scale)
        out = np.zeros((width * height, ch), np.float32)
        self._chk(lib.shb_downsample(self.h, width, height, scale, ch, _f(img), _f(out)))
        return out

    def downsample_mask(self, mask, width, height, scale):
        mask = np.ascontiguousarray(mask, np.uint8)
        out = np.zeros(width * height, np.uint8)
        self._chk(lib.shb_downsample_mask(self.h, width, height, scale, _u8(mask), _u8(out)))
        return out

    def dilate(self, data, mask, width, height, passes):
        data = np.array(data, np.float32, order="C", copy=True)
        mask = np.array(mask, np.uint8, order="C", copy=True)
        ch = data.size // max(1, width * height)
        self._chk(lib.shb_dilate(self.h, width, height, ch, _f(data), _u8(mask), passes))
        return data, mask

    # -- rays
    def trace_rays(self, rays, any_hit=False, counts=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = rays.size
        hits = np.zeros(n, HIT_DTYPE)
        nv = np.zeros(n, np.uint32) if counts else None
        tt = np.zeros(n, np.uint32) if counts else None
        p32 = C.POINTER(C.c_uint32)
        self._chk(lib.shb_trace_rays(
            self.h, rays.ctypes.data_as(C.c_void_p), n, 1 if any_hit else 0,
            hits.ctypes.data_as(C.c_void_p),
            nv.ctypes.data_as(p32) if counts else None,
            tt.ctypes.data_as(p32) if counts else None))
        return (hits, nv, tt) if counts else hits

    def bvh_export(self):
        """The flattened BVH of the uploaded scene as the kernels see it:
        (nodes: (n, 80) uint8, tris: (m, 12) float32 = 3 float4 per
        triangle in leaf order, v0.w holding the triangle id's bits)."""
        nn, nt = C.c_uint32(0), C.c_uint32(0)
        self._chk(lib.shb_bvh_export(self.h, None, C.byref(nn), None, C.byref(nt)))
        nodes = np.zeros((nn.value, 80), np.uint8)
        tris = np.zeros((nt.value, 12), np.float32)
        self._chk(lib.shb_bvh_export(self.h, nodes.ctypes.data_as(C.c_void_p), C.byref(nn),
                                     tris.ctypes.data_as(C.c_void_p), C.byref(nt)))
        return nodes, tris

    # -- tuning / params
    def tuning(self):
        t = ShbTuning()
        self._chk(lib.shb_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = ShbTuning(**cur)
        self._chk(lib.shb_set_tuning(self.h, C.byref(t)))

    def bake_adjoint(self, samples, bounces, dloss_dsh, seed=12345, flags=0,
                     num_light_samples=1):
        """Path-replay adjoint of bake(): given dL/dSH at the supersampled
        resolution ((n, 27) like bake's sh), returns dL/d(light_scale)
        (lights, 3), dL/d(albedo_scale) (materials, 3) and the stats."""
        n = self.width * self.ss * self.height * self.ss
        g = np.ascontiguousarray(dloss_dsh, np.float32).reshape(n, 27)
        nl, nm = len(self.scene.lights), len(self.scene.materials)
        dl = np.zeros((nl, 3), np.float32)
        dm = np.zeros((nm, 3), np.float32)
        st = ShbBakeStats()
        cfg = self.cfg(samples, bounces, seed, flags, num_light_samples)
        self._chk(lib.shb_bake_adjoint(self.h, C.byref(cfg), _f(g), _f(dl), _f(dm), C.byref(st)))
        return dl, dm, st.as_dict()

    def env_sh_latlong(self, pixels=None, shape=None):
        """SH9 (27,) of an (h, w, ch >= 3) equirectangular float map on the
        device; pixels None reuses the map of the previous call (shape =
        (h, w, ch) then). Returns (sh, kernel seconds)."""
        if pixels is not None:
            px = np.ascontiguousarray(pixels, np.float32)
            h, w, ch = px.shape
            ptr = _f(px)
        else:
            h, w, ch = shape
            ptr = None
        out = np.zeros(27, np.float32)
        sec = C.c_double(0.0)
        self._chk(lib.shb_env_sh_latlong(self.h, w, h, ch, ptr, _f(out), C.byref(sec)))
        return out, sec.value

    def env_sh_latlong_noalloc(self, pixels_device, width, height, channels, sh_out_device):
        """shb_env_sh_latlong_noalloc: pixels_device and sh_out_device are
        device addresses (int) of width*height*channels and 27 floats."""
        self._chk(lib.shb_env_sh_latlong_noalloc(self.h, width, height, channels,
                                                 C.c_void_p(pixels_device),
                                                 C.c_void_p(sh_out_device)))

    def set_params(self, light_scale=None, albedo_scale=None):
        ls = None if light_scale is None else np.ascontiguousarray(light_scale, np.float32)
        al = None if albedo_scale is None else np.ascontiguousarray(albedo_scale, np.float32)
        self._chk(lib.shb_set_params(self.h, _f(ls), _f(al)))


def bake_batched(bakers, samples, bounces, seeds=None, flags=0, want=True,
                 num_light_samples=1):
    """Bakes several Bakers (each with its own scene and surface, all on one
    device) in one batched run (shared wavefront batches, hipGraph replay).
    samples / bounces / seeds / num_light_samples: one value per job or a
    scalar. Returns ([sh], [vis], stats) like Baker.bake per job."""
    n = len(bakers)
    as_list = lambda v: list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v] * n
    samples, bounces = as_list(samples), as_list(bounces)
    seeds = as_list(12345 if seeds is None else seeds)
    nls = as_list(num_light_samples)
    cfgs = (ShbBakeCfg * n)(*[Baker.cfg(int(samples[i]), int(bounces[i]), int(seeds[i]), flags,
                                        int(nls[i])) for i in range(n)])
    ctxs = (C.c_void_p * n)(*[b.h for b in bakers])
    shs, viss = [], []
    for b in bakers:
        m = b.width * b.ss * b.height * b.ss
        shs.append(np.zeros((m, 27), np.float32) if want else None)
        viss.append(np.zeros(m, np.float32) if want else None)
    sh_p = (_fp * n)(*[_f(a) for a in shs]) if want else None
    vis_p = (_fp * n)(*[_f(a) for a in viss]) if want else None
    st = ShbBakeStats()
    rc = lib.shb_bake_batched(ctxs, n, cfgs, sh_p, vis_p, C.byref(st))
    if rc != SHB_OK:
        raise ShbError(rc, lib.shb_last_error(bakers[0].h).decode())
    return shs, viss, st.as_dict()


def env_sh_preetham(sun_direction):
    """SH9 (27,) of the clamped Preetham clear sky for a sun direction
    (host computation)."""
    get_lib()
    out = np.zeros(27, np.float32)
    rc = lib.shb_env_sh_preetham(_f(np.ascontiguousarray(sun_direction, np.float32)), _f(out))
    if rc != SHB_OK:
        raise ShbError(rc, "bad sun direction")
    return out


def save_lightmap(path, sh, vis, width, height, mode=SAVE_COMBINED):
    """Writes an SH lightmap ((width*height, 27) and (width*height,)) as
    OpenEXR (HALF, ZIP) in one of the reference's three layouts."""
    get_lib()
    sh = np.ascontiguousarray(sh, np.float32).reshape(width * height, 27)
    vis = None if vis is None else np.ascontiguousarray(vis, np.float32).reshape(width * height)
    err = C.create_string_buffer(512)
    rc = lib.shb_save_lightmap(str(path).encode(), width, height, _f(sh), _f(vis), mode, err, 512)
    if rc != SHB_OK:
        raise ShbError(rc, err.value.decode())


def load_gltf(path):
    """Loads a .gltf / .glb with the library's loader. Returns (scene, env):
    a shb.scenes.Scene (flat arrays, materials carry 'name' and
    'albedo_path') and a dict describing the environment ('type' ENV_*,
    'pixels' (h, w, 3) for a skybox texture, 'sun_direction',
    'sun_intensity', 'path')."""
    from . import scenes
    get_lib()
    h = C.c_void_p()
    err = C.create_string_buffer(1024)
    rc = lib.shb_load_gltf(str(path).encode(), C.byref(h), err, 1024)
    if rc != SHB_OK:
        raise ShbError(rc, err.value.decode(errors="replace"))
    try:
        s = _scene_from_desc(lib.shb_gltf_scene(h).contents, h)
        e = ShbEnvironmentDesc()
        lib.shb_gltf_environment(h, C.byref(e))
        env = {"type": e.type, "sun_direction": tuple(e.sun_direction),
               "sun_intensity": e.sun_intensity,
               "path": e.path.decode() if e.path else None, "pixels": None}
        if e.type == ENV_TEXTURE and e.pixels:
            env["pixels"] = np.ctypeslib.as_array(
                e.pixels, shape=(e.width * e.height * 3,)).copy().reshape(e.height, e.width, 3)
        return s, env
    finally:
        lib.shb_gltf_free(h)


def _scene_from_desc(d, gltf_handle=None):
    """A shb.scenes.Scene copied from a shb_scene_desc."""
    from . import scenes
    nv, ni = d.num_vertices, d.num_indices

    def arr(ptr, n, dt, shape):
        if n == 0 or not ptr:
            return np.zeros(shape, dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).reshape(shape)

    if True:
        h = gltf_handle
        s = scenes.Scene()
        s.positions = arr(d.positions, nv * 3, np.float32, (nv, 3))
        s.normals = arr(d.normals, nv * 3, np.float32, (nv, 3))
        s.texture_uvs = arr(d.texture_uvs, nv * 2, np.float32, (nv, 2))
        s.lightmap_uvs = arr(d.lightmap_uvs, nv * 2, np.float32, (nv, 2))
        s.tangents = arr(d.tangents, nv * 4, np.float32, (nv, 4))
        s.indices = arr(d.indices, ni, np.uint32, (ni,))
        s.geometries = []
        for i in range(d.num_geometries):
            g = d.geometries[i]
            s.geometries.append(dict(vertex_offset=g.vertex_offset, vertex_count=g.vertex_count,
                                     index_offset=g.index_offset, index_count=g.index_count,
                                     material_id=g.material_id, flags=g.flags,
                                     transform=tuple(g.transform)))
        s.materials = []
        for i in range(d.num_materials):
            m = d.materials[i]
            name = lib.shb_gltf_material_name(h, i) if h else None
            ap = lib.shb_gltf_albedo_path(h, i) if h else None
            s.materials.append(dict(width=m.albedo_width, height=m.albedo_height,
                                    channels=m.albedo_channels, offset=m.albedo_offset,
                                    emission=m.emission_intensity,
                                    name=name.decode() if name is not None else "",
                                    albedo_path=ap.decode() if ap is not None else None))
        s.texels = arr(d.texel_pool, d.texel_pool_size, np.uint8, (d.texel_pool_size,))
        s.lights = []
        for i in range(d.num_lights):
            l = d.lights[i]
            s.lights.append(dict(type=l.type, position=tuple(l.position),
                                 direction=tuple(l.direction), color=tuple(l.color),
                                 intensity=l.intensity, cos_inner_cone=l.cos_inner_cone,
                                 cos_outer_cone=l.cos_outer_cone,
                                 geometry_index=l.geometry_index,
                                 material_index=l.material_index))
        return s


def geometry_scales(scene, density_multiplier=1.0):
    """Per-geometry atlas scales (CalculateGeometryScales)."""
    get_lib()
    desc, keep = scene.to_desc()
    out = np.zeros(max(len(scene.geometries), 1), np.float32)
    rc = lib.shb_atlas_geometry_scales(C.byref(desc), density_multiplier, _f(out))
    if rc != SHB_OK:
        raise ShbError(rc, "bad scene for atlas scales")
    del keep
    return out[:len(scene.geometries)]


def create_atlas(scene, resolution, padding, density_multiplier=1.0):
    """Lightmap atlas for `scene`: returns (scene with lightmap UVs and
    split vertices, info dict with width, height, charts, texels_per_unit,
    and per output vertex xref (the geometry-local source vertex) and
    chart)."""
    get_lib()
    desc, keep = scene.to_desc()
    h = C.c_void_p()
    err = C.create_string_buffer(1024)
    rc = lib.shb_atlas_create(C.byref(desc), resolution, padding, density_multiplier,
                              C.byref(h), err, 1024)
    del keep
    if rc != SHB_OK:
        raise ShbError(rc, err.value.decode(errors="replace"))
    try:
        out = _scene_from_desc(lib.shb_atlas_scene(h).contents)
        for m_out, m_in in zip(out.materials, scene.materials):
            for k in ("name", "albedo_path"):
                if k in m_in:
                    m_out[k] = m_in[k]
        w, hh, nch = C.c_int32(), C.c_int32(), C.c_uint32()
        tpu = C.c_float()
        xr = C.POINTER(C.c_uint32)()
        co = C.POINTER(C.c_uint32)()
        lib.shb_atlas_info(h, C.byref(w), C.byref(hh), C.byref(nch), C.byref(tpu), C.byref(xr),
                           C.byref(co))
        nvo = out.positions.shape[0]
        xref = np.ctypeslib.as_array(xr, shape=(max(nvo, 1),)).copy()[:nvo]
        chart = np.ctypeslib.as_array(co, shape=(max(nvo, 1),)).copy()[:nvo]
        return out, {"width": w.value, "height": hh.value, "charts": nch.value,
                     "texels_per_unit": tpu.value, "xref": xref, "chart": chart}
    finally:
        lib.shb_atlas_free(h)
