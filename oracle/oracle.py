"""ORACLE bindings — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU restatement of the
reference bake path (see oracle.cpp). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module, always as the checker
(or the timed CPU baseline), never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_REF = 0   # std::mt19937 + libstdc++ distributions (reference stream)
MODE_GPU = 1   # Philox4x32-10 / xoshiro128** stream of the HIP kernels


class OrcBakeStats(C.Structure):
    _fields_ = [("valid_texels", C.c_uint64), ("paths", C.c_uint64),
                ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("seconds", C.c_double), ("threads", C.c_int)]


def build():
    """Compiles liboracle.so (g++, a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        u32p = C.POINTER(C.c_uint32)
        u8p = C.POINTER(C.c_uint8)
        L.orc_scene_create.restype = vp
        L.orc_scene_create.argtypes = [vp, C.c_int]
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_trace_rays.argtypes = [vp, vp, C.c_uint32, C.c_int, vp, u32p, u32p]
        L.orc_cwbvh_trace.argtypes = [vp, C.c_uint32, vp, vp, C.c_uint32, C.c_int, vp,
                                      u32p, u32p]
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_xoshiro128ss.argtypes = [u32p, C.c_int, u32p]
        L.orc_site_uniforms.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, fp]
        L.orc_sh_accumulate.argtypes = [fp, fp, fp]
        L.orc_env_latlong.argtypes = [fp, C.c_int, C.c_int, C.c_int, fp,
                                      C.POINTER(C.c_double)]
        L.orc_env_preetham.argtypes = [fp, fp]
        L.orc_set_fault.argtypes = [C.c_int]
        L.orc_srgb_to_linear.argtypes = [C.c_uint8]
        L.orc_srgb_to_linear.restype = C.c_float
        L.orc_sample_material.argtypes = [fp, C.c_float, C.c_float, fp]
        L.orc_material_eval.argtypes = [vp, C.c_int, fp, fp]
        L.orc_light_radiance.argtypes = [vp, C.c_int, fp, fp, fp, C.c_int, fp, C.c_int, fp, fp]
        L.orc_sample_area_light.argtypes = [vp, C.c_int, C.c_uint32, C.c_float, C.c_float, fp]
        L.orc_evaluate_light_samples.argtypes = [vp, fp, fp, fp, C.c_int, fp, C.c_uint,
                                                 C.c_uint32, fp]
        L.orc_rasterize.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
        L.orc_validity_mask.argtypes = [vp, C.c_size_t, u8p]
        L.orc_downsample_mask.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p]
        L.orc_downsample.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, fp]
        L.orc_dilate.argtypes = [C.c_int, C.c_int, C.c_int, fp, u8p, C.c_int]
        L.orc_bake.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_int,
                               C.c_uint64, C.c_uint32, C.c_int, C.c_int, fp, fp, fp,
                               C.POINTER(OrcBakeStats)]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _vec(x, n):
    return np.ascontiguousarray(np.asarray(x, np.float32).reshape(n))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    p = C.POINTER(C.c_uint32)
    lib().orc_philox4x32_10(c.ctypes.data_as(p), k.ctypes.data_as(p), o.ctypes.data_as(p))
    return o


def xoshiro(state, n):
    s = np.ascontiguousarray(state, np.uint32)
    o = np.zeros(n, np.uint32)
    p = C.POINTER(C.c_uint32)
    lib().orc_xoshiro128ss(s.ctypes.data_as(p), n, o.ctypes.data_as(p))
    return o


def site_uniforms(seed, texel, sample, depth, branch, kind, job, n):
    o = np.zeros(n, np.float32)
    lib().orc_site_uniforms(seed, texel, sample, depth, branch, kind, job, n, _f(o))
    return o


def sh_accumulate(radiance, direction, sh=None):
    sh = np.zeros(27, np.float32) if sh is None else np.ascontiguousarray(sh, np.float32)
    lib().orc_sh_accumulate(_f(_vec(radiance, 3)), _f(_vec(direction, 3)), _f(sh))
    return sh


def env_latlong(pixels):
    """ProjectLatLongMap of an (h, w, ch) float map: (reference serial float
    sum, FP64 sum of the same terms), each (27,)."""
    px = np.ascontiguousarray(pixels, np.float32)
    h, w, ch = px.shape
    ref = np.zeros(27, np.float32)
    s64 = np.zeros(27, np.float64)
    lib().orc_env_latlong(_f(px), w, h, ch, _f(ref), s64.ctypes.data_as(C.POINTER(C.c_double)))
    return ref, s64


def env_preetham(sun):
    out = np.zeros(27, np.float32)
    lib().orc_env_preetham(_f(_vec(sun, 3)), _f(out))
    return out


FAULT_CONT_NO_COSINE, FAULT_NEE_NO_COSINE, FAULT_SPOT_NO_FALLOFF, FAULT_AREA_PDF_TOTAL = 1, 2, 3, 4


def set_fault(fault):
    """Plants a known mistake in the oracle (tests that show a check catches
    it); 0 removes it."""
    lib().orc_set_fault(fault)


def srgb_to_linear(x):
    return lib().orc_srgb_to_linear(x)


def sample_material(n, r1, r2):
    o = np.zeros(4, np.float32)
    lib().orc_sample_material(_f(_vec(n, 3)), r1, r2, _f(o))
    return o[:3], o[3]


def validity_mask(points):
    points = np.ascontiguousarray(points)
    m = np.zeros(points.size, np.uint8)
    lib().orc_validity_mask(points.ctypes.data_as(C.c_void_p), points.size, _u8(m))
    return m


def downsample_mask(mask, width, height, scale):
    mask = np.ascontiguousarray(mask, np.uint8)
    out = np.zeros(width * height, np.uint8)
    lib().orc_downsample_mask(_u8(mask), width, height, scale, _u8(out))
    return out


def downsample(img, in_w, in_h, scale):
    img = np.ascontiguousarray(img, np.float32)
    ch = img.size // (in_w * in_h)
    out = np.zeros(((in_w // scale) * (in_h // scale), ch), np.float32)
    lib().orc_downsample(_f(img), in_w, in_h, scale, ch, _f(out))
    return out


def dilate(data, mask, width, height, passes):
    data = np.array(data, np.float32, order="C", copy=True)
    mask = np.array(mask, np.uint8, order="C", copy=True)
    ch = data.size // max(1, width * height)
    lib().orc_dilate(width, height, ch, _f(data), _u8(mask), passes)
    return data, mask


class Oracle:
    """The oracle's copy of a scene (flat shb_scene_desc arrays)."""

    def __init__(self, scene, bvh=True):
        desc, keep = scene.to_desc()
        self.h = lib().orc_scene_create(C.byref(desc), 1 if bvh else 0)
        del keep
        self.scene = scene

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def trace(self, rays, mode=0, counts=False):
        """mode 0: FP32 watertight closest hit; 1: FP64 brute force;
        2: FP32 any hit."""
        from shb import HIT_DTYPE, RAY_DTYPE
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(rays.size, HIT_DTYPE)
        nv = np.zeros(rays.size, np.uint32)
        tt = np.zeros(rays.size, np.uint32)
        p = C.POINTER(C.c_uint32)
        lib().orc_trace_rays(self.h, rays.ctypes.data_as(C.c_void_p), rays.size, mode,
                             hits.ctypes.data_as(C.c_void_p), nv.ctypes.data_as(p),
                             tt.ctypes.data_as(p))
        return (hits, nv, tt) if counts else hits

    def rasterize(self, width, height, ss=1):
        from shb import SURFACE_DTYPE
        out = np.zeros(width * ss * height * ss, SURFACE_DTYPE)
        lib().orc_rasterize(self.h, width, height, ss, out.ctypes.data_as(C.c_void_p))
        return out

    def material(self, mat, uv):
        o = np.zeros(7, np.float32)
        lib().orc_material_eval(self.h, mat, _f(_vec(uv, 2)), _f(o))
        return o[:3], o[3], o[4:]

    def light_radiance(self, li, P, N, wo=(0, 0, 1), mat=-1, uv=(0, 0), with_brdf=0,
                       area=(0, 0, 0)):
        o = np.zeros(11, np.float32)
        lib().orc_light_radiance(self.h, li, _f(_vec(P, 3)), _f(_vec(N, 3)), _f(_vec(wo, 3)),
                                 mat, _f(_vec(uv, 2)), with_brdf, _f(_vec(area, 3)), _f(o))
        return o[:3], o[3:6], o[6:9], o[9], o[10]

    def sample_area_light(self, li, tri, u1, u2):
        o = np.zeros(10, np.float32)
        lib().orc_sample_area_light(self.h, li, tri, u1, u2, _f(o))
        return o[:3], o[3:6], o[6:9], o[9]

    def evaluate_light_samples(self, P, N, wo, mat, uv, n, seed):
        o = np.zeros(3, np.float32)
        lib().orc_evaluate_light_samples(self.h, _f(_vec(P, 3)), _f(_vec(N, 3)), _f(_vec(wo, 3)),
                                         mat, _f(_vec(uv, 2)), n, seed, _f(o))
        return o

    def bake(self, points, width, height, samples, bounces, mode=MODE_REF, seed=12345, job=0,
             threads=0, row_stride=1, num_light_samples=1):
        """BakeSHLightMap over (width*ss x height*ss) points given as the
        full supersampled size (width, height here are that size)."""
        points = np.ascontiguousarray(points)
        n = width * height
        assert points.size == n
        sh = np.zeros((n, 27), np.float32)
        vis = np.zeros(n, np.float32)
        var = np.zeros((n, 27), np.float32)
        st = OrcBakeStats()
        lib().orc_bake(self.h, points.ctypes.data_as(C.c_void_p), width, height, samples, bounces,
                       num_light_samples, mode, seed, job, threads, row_stride, _f(sh), _f(vis),
                       _f(var), C.byref(st))
        stats = {k: getattr(st, k) for k, _ in st._fields_}
        return sh, vis, var, stats


def cwbvh_trace(nodes, tris, rays, any_hit=False):
    """Traversal of an exported compressed 8-wide BVH (nodes (n, 80) uint8,
    tris (m, 12) float32) over a ray list: (hits, node visits, triangle
    tests) per ray, recounted on the host."""
    from shb import HIT_DTYPE, RAY_DTYPE
    nodes = np.ascontiguousarray(nodes, np.uint8)
    tris = np.ascontiguousarray(tris, np.float32)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(rays.size, HIT_DTYPE)
    nv = np.zeros(rays.size, np.uint32)
    tt = np.zeros(rays.size, np.uint32)
    p = C.POINTER(C.c_uint32)
    lib().orc_cwbvh_trace(nodes.ctypes.data_as(C.c_void_p), nodes.shape[0],
                          tris.ctypes.data_as(C.c_void_p), rays.ctypes.data_as(C.c_void_p),
                          rays.size, 1 if any_hit else 0, hits.ctypes.data_as(C.c_void_p),
                          nv.ctypes.data_as(p), tt.ctypes.data_as(p))
    return hits, nv, tt
