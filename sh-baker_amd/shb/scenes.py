"""Scene building, a per-chart shelf atlas packer and the seeded synthetic
benchmark scenes.

The reference reads glTF and re-UVs with xatlas (src/loader.cpp,
src/atlas.cpp); the benchmark workloads here are generated from a seed
instead: every face (or cylinder side / cap) is its own lightmap chart,
packed area-proportionally by a shelf packer with `padding` texels around
each chart.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import (GEOM_HAS_LIGHTMAP_UVS, GEOM_HAS_NORMALS, GEOM_HAS_TANGENTS,
               GEOM_HAS_TEXTURE_UVS, LIGHT_AREA, LIGHT_DIRECTIONAL, LIGHT_POINT, LIGHT_SPOT,
               ShbGeometry,
               ShbLight, ShbMaterial, ShbSceneDesc)

IDENTITY = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)


def linear_to_srgb8(x: float) -> int:
    """src/colorspace.h:16-21 LinearToSRGB (used for baseColorFactor)."""
    x = np.float32(x)
    if x <= np.float32(0.0031308):
        return int(np.rint(np.float32(x * np.float32(12.92) * np.float32(255.0))))
    v = np.float32(1.055) * np.float32(np.power(x, np.float32(1.0 / 2.4), dtype=np.float32)) \
        - np.float32(0.055)
    return int(np.rint(np.float32(v * np.float32(255.0))))


class Scene:
    """Flat scene arrays in the layout of shb_scene_desc."""

    def __init__(self):
        self.positions = np.zeros((0, 3), np.float32)
        self.normals = np.zeros((0, 3), np.float32)
        self.texture_uvs = np.zeros((0, 2), np.float32)
        self.lightmap_uvs = np.zeros((0, 2), np.float32)
        self.tangents = np.zeros((0, 4), np.float32)
        self.indices = np.zeros(0, np.uint32)
        self.geometries = []   # dicts
        self.materials = []    # dicts
        self.texels = np.zeros(0, np.uint8)
        self.lights = []       # dicts

    @property
    def num_triangles(self):
        return sum(g["index_count"] // 3 for g in self.geometries)

    def to_desc(self):
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        pos = arr(self.positions, np.float32)
        nrm = arr(self.normals, np.float32)
        tuv = arr(self.texture_uvs, np.float32)
        luv = arr(self.lightmap_uvs, np.float32)
        tan = arr(self.tangents, np.float32)
        idx = arr(self.indices, np.uint32)
        tex = arr(self.texels if self.texels.size else np.zeros(1, np.uint8), np.uint8)
        ng, nm, nl = len(self.geometries), len(self.materials), len(self.lights)
        geos = (ShbGeometry * max(ng, 1))()
        for i, g in enumerate(self.geometries):
            geos[i].vertex_offset = g["vertex_offset"]
            geos[i].vertex_count = g["vertex_count"]
            geos[i].index_offset = g["index_offset"]
            geos[i].index_count = g["index_count"]
            geos[i].material_id = g["material_id"]
            geos[i].flags = g["flags"]
            for k in range(12):
                geos[i].transform[k] = g["transform"][k]
        mats = (ShbMaterial * max(nm, 1))()
        for i, m in enumerate(self.materials):
            mats[i].albedo_width = m["width"]
            mats[i].albedo_height = m["height"]
            mats[i].albedo_channels = m["channels"]
            mats[i].albedo_offset = m["offset"]
            mats[i].emission_intensity = m["emission"]
        lts = (ShbLight * max(nl, 1))()
        for i, l in enumerate(self.lights):
            lts[i].type = l["type"]
            for k in range(3):
                lts[i].position[k] = l.get("position", (0, 0, 0))[k]
                lts[i].direction[k] = l.get("direction", (0, 0, -1))[k]
                lts[i].color[k] = l.get("color", (1, 1, 1))[k]
            lts[i].intensity = l.get("intensity", 1.0)
            lts[i].cos_inner_cone = l.get("cos_inner_cone", 1.0)
            lts[i].cos_outer_cone = l.get("cos_outer_cone", 0.70710678118654752440)
            lts[i].geometry_index = l.get("geometry_index", -1)
            lts[i].material_index = l.get("material_index", -1)
        keep += [geos, mats, lts]
        fp = C.POINTER(C.c_float)
        d = ShbSceneDesc()
        d.num_vertices = pos.shape[0]
        d.positions = pos.ctypes.data_as(fp)
        d.normals = nrm.ctypes.data_as(fp) if nrm.size else None
        d.texture_uvs = tuv.ctypes.data_as(fp) if tuv.size else None
        d.lightmap_uvs = luv.ctypes.data_as(fp) if luv.size else None
        d.tangents = tan.ctypes.data_as(fp) if tan.size else None
        d.num_indices = idx.size
        d.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
        d.num_geometries = ng
        d.geometries = geos
        d.num_materials = nm
        d.materials = mats
        d.texel_pool_size = self.texels.size
        d.texel_pool = tex.ctypes.data_as(C.POINTER(C.c_uint8))
        d.num_lights = nl
        d.lights = lts
        keep.append(d)
        return d, keep


class SceneBuilder:
    """Accumulates geometries with per-chart local lightmap coordinates (in
    world units), then packs all charts into one atlas."""

    def __init__(self):
        self.pos, self.nrm, self.tuv, self.tan, self.idx = [], [], [], [], []
        self.local = []        # per-vertex chart-local coords (world units)
        self.chart_of = []     # per-vertex chart id
        self.charts = []       # (w, h) extents
        self.geoms = []
        self.materials = []
        self.texels = bytearray()
        self.lights = []
        self.nv = 0
        self.ni = 0

    def add_material(self, rgba=(255, 255, 255, 255), emission=0.0, pixels=None):
        """1x1 RGBA sRGB8 albedo (the loader's baseColorFactor form,
        src/loader.cpp:487-505), or a (h, w, c) uint8 texture, or None for
        no texture (albedo 0.8)."""
        if pixels is None and rgba is None:
            self.materials.append(dict(width=0, height=0, channels=0, offset=0,
                                       emission=float(emission)))
            return len(self.materials) - 1
        if pixels is None:
            pixels = np.array(rgba, np.uint8).reshape(1, 1, len(rgba))
        pixels = np.ascontiguousarray(pixels, np.uint8)
        h, w, c = pixels.shape
        off = len(self.texels)
        self.texels += pixels.tobytes()
        self.materials.append(dict(width=w, height=h, channels=c, offset=off,
                                   emission=float(emission)))
        return len(self.materials) - 1

    def new_chart(self, local):
        local = np.asarray(local, np.float64)
        lo = local.min(axis=0)
        ext = local.max(axis=0) - lo
        self.charts.append((float(ext[0]), float(ext[1])))
        return len(self.charts) - 1, local - lo

    def add_geometry(self, pos, nrm, tan, tri, chart_ids, local, material, tuv=None,
                     transform=IDENTITY):
        pos = np.asarray(pos, np.float32)
        n = pos.shape[0]
        self.pos.append(pos)
        self.nrm.append(np.asarray(nrm, np.float32))
        self.tan.append(np.asarray(tan, np.float32))
        self.tuv.append(np.zeros((n, 2), np.float32) if tuv is None else np.asarray(tuv, np.float32))
        self.local.append(np.asarray(local, np.float64))
        self.chart_of.append(np.asarray(chart_ids, np.int64))
        tri = np.asarray(tri, np.uint32).reshape(-1)
        self.idx.append(tri)
        flags = GEOM_HAS_NORMALS | GEOM_HAS_LIGHTMAP_UVS | GEOM_HAS_TANGENTS
        if tuv is not None:
            flags |= GEOM_HAS_TEXTURE_UVS
        self.geoms.append(dict(vertex_offset=self.nv, vertex_count=n, index_offset=self.ni,
                               index_count=tri.size, material_id=material, flags=flags,
                               transform=tuple(transform)))
        self.nv += n
        self.ni += tri.size
        return len(self.geoms) - 1

    # -------------------------------------------------------------- shapes
    def add_quads(self, faces, material, grid=None):
        """Planar parallelogram faces (p0, e1, e2), normal along e1 x e2, each
        its own chart; grid[i] = (nu, nv) subdivisions."""
        P, N, T, TRI, CH, LOC = [], [], [], [], [], []
        base = 0
        for fi, (p0, e1, e2) in enumerate(faces):
            p0, e1, e2 = (np.asarray(x, np.float64) for x in (p0, e1, e2))
            nu, nvv = (1, 1) if grid is None else grid[fi]
            nrm = np.cross(e1, e2)
            nrm /= np.linalg.norm(nrm)
            t = e1 / np.linalg.norm(e1)
            uu, vv = np.meshgrid(np.arange(nu + 1) / nu, np.arange(nvv + 1) / nvv, indexing="xy")
            uu, vv = uu.ravel(), vv.ravel()
            p = p0[None] + uu[:, None] * e1[None] + vv[:, None] * e2[None]
            loc = np.stack([uu * np.linalg.norm(e1), vv * np.linalg.norm(e2)], 1)
            cid, loc = self.new_chart(loc)
            for j in range(nvv):
                for i in range(nu):
                    a = base + j * (nu + 1) + i
                    b, c, d = a + 1, a + (nu + 1) + 1, a + (nu + 1)
                    TRI += [a, b, c, a, c, d]
            P.append(p)
            N.append(np.repeat(nrm[None], len(p), 0))
            T.append(np.repeat(np.append(t, 1.0)[None], len(p), 0))
            CH.append(np.full(len(p), cid))
            LOC.append(loc)
            base += len(p)
        return self.add_geometry(np.concatenate(P), np.concatenate(N), np.concatenate(T), TRI,
                                 np.concatenate(CH), np.concatenate(LOC), material)

    def add_box(self, center, half, R=np.eye(3), material=0, inward=False, cell=None):
        """Six faces of an oriented box; outward normals, or inward for a
        room. With `cell`, each face is subdivided into cells of about that
        size (one chart per face either way)."""
        center, half = np.asarray(center, np.float64), np.asarray(half, np.float64)
        faces, grid = [], []
        for a in range(3):
            b, c = (a + 1) % 3, (a + 2) % 3
            ax, bx, cx = R[:, a], R[:, b], R[:, c]
            for s in (1, -1):
                p0 = center + s * half[a] * ax - half[b] * bx - half[c] * cx
                eb, ec = 2 * half[b] * bx, 2 * half[c] * cx
                # bx x cx = ax: (eb, ec) faces +ax, (ec, eb) faces -ax
                e1, e2 = (eb, ec) if (s > 0) != inward else (ec, eb)
                faces.append((p0, e1, e2))
                if cell:
                    grid.append((max(1, round(np.linalg.norm(e1) / cell)),
                                 max(1, round(np.linalg.norm(e2) / cell))))
        return self.add_quads(faces, material, grid if cell else None)

    def add_boxes(self, centers, halves, Rs, materials):
        """add_box for many unsubdivided outward boxes at once: the same
        vertices, normals, tangents, charts and triangles, bit for bit (the
        lengths are taken with the per-vector dot that np.linalg.norm uses),
        with the arithmetic done over all faces together."""
        centers = np.asarray(centers, np.float64)
        halves = np.asarray(halves, np.float64)
        Rs = np.asarray(Rs, np.float64)
        n = len(centers)
        P0, E1, E2 = [], [], []
        for a in range(3):
            b, c = (a + 1) % 3, (a + 2) % 3
            ax, bx, cx = Rs[:, :, a], Rs[:, :, b], Rs[:, :, c]
            for sgn in (1, -1):
                p0 = (centers + (sgn * halves[:, a])[:, None] * ax - halves[:, b][:, None] * bx
                      - halves[:, c][:, None] * cx)
                eb, ec = (2 * halves[:, b])[:, None] * bx, (2 * halves[:, c])[:, None] * cx
                e1, e2 = (eb, ec) if sgn > 0 else (ec, eb)
                P0.append(p0)
                E1.append(e1)
                E2.append(e2)
        P0, E1, E2 = (np.stack(v, 1) for v in (P0, E1, E2))  # (n, 6, 3)

        def lengths(v):
            f = v.reshape(-1, 3)
            return np.sqrt(np.array([x.dot(x) for x in f])).reshape(v.shape[:-1])

        nrm = np.cross(E1, E2)
        nrm /= lengths(nrm)[..., None]
        l1, l2 = lengths(E1), lengths(E2)
        t = E1 / l1[..., None]
        uu = np.array([0.0, 1.0, 0.0, 1.0])
        vv = np.array([0.0, 0.0, 1.0, 1.0])
        pos = P0[:, :, None] + uu[None, None, :, None] * E1[:, :, None] + \
            vv[None, None, :, None] * E2[:, :, None]  # (n, 6, 4, 3)
        loc = np.stack([uu[None, None] * l1[..., None], vv[None, None] * l2[..., None]], -1)
        tri = []
        for fi in range(6):
            a4 = 4 * fi
            tri += [a4, a4 + 1, a4 + 3, a4, a4 + 3, a4 + 2]
        nrm4 = np.repeat(nrm[:, :, None], 4, 2)
        tan4 = np.concatenate([np.repeat(t[:, :, None], 4, 2), np.ones((n, 6, 4, 1))], -1)
        first = []
        for k in range(n):
            base_chart = len(self.charts)
            for fi in range(6):
                self.charts.append((float(l1[k, fi]), float(l2[k, fi])))
            ch = np.repeat(np.arange(base_chart, base_chart + 6), 4)
            first.append(self.add_geometry(pos[k].reshape(24, 3), nrm4[k].reshape(24, 3),
                                           tan4[k].reshape(24, 4), tri, ch,
                                           loc[k].reshape(24, 2), materials[k]))
        return first

    def add_cylinder(self, cx, cz, r, y0, y1, segments, rings, material):
        th = np.arange(segments + 1) / segments * 2 * math.pi
        ys = y0 + (y1 - y0) * np.arange(rings + 1) / rings
        TH, Y = np.meshgrid(th, ys, indexing="xy")
        TH, Y = TH.ravel(), Y.ravel()
        side = np.stack([cx + r * np.cos(TH), Y, cz + r * np.sin(TH)], 1)
        sn = np.stack([np.cos(TH), np.zeros_like(TH), np.sin(TH)], 1)
        st = np.stack([-np.sin(TH), np.zeros_like(TH), np.cos(TH), np.ones_like(TH)], 1)
        cid, sloc = self.new_chart(np.stack([r * TH, Y], 1))
        tri = []
        w = segments + 1
        for i in range(rings):
            for j in range(segments):
                v00, v01 = i * w + j, i * w + j + 1
                v10, v11 = v00 + w, v01 + w
                tri += [v00, v10, v11, v00, v11, v01]
        P, N, T, CH, LOC = [side], [sn], [st], [np.full(len(side), cid)], [sloc]
        base = len(side)
        ring_th = np.arange(segments) / segments * 2 * math.pi
        for top in (True, False):
            yy = y1 if top else y0
            pts = np.stack([cx + r * np.cos(ring_th), np.full(segments, yy), cz + r * np.sin(ring_th)], 1)
            nrm = np.array([0, 1 if top else -1, 0], np.float64)
            ccid, cloc = self.new_chart(np.stack([pts[:, 0], pts[:, 2]], 1))
            for k in range(1, segments - 1):
                tri += ([base, base + k + 1, base + k] if top else [base, base + k, base + k + 1])
            P.append(pts)
            N.append(np.repeat(nrm[None], segments, 0))
            T.append(np.repeat(np.array([[1, 0, 0, 1.0]]), segments, 0))
            CH.append(np.full(segments, ccid))
            LOC.append(cloc)
            base += segments
        return self.add_geometry(np.concatenate(P), np.concatenate(N), np.concatenate(T), tri,
                                 np.concatenate(CH), np.concatenate(LOC), material)

    def add_triangle(self, p, material):
        p = np.asarray(p, np.float64)
        e1, e2 = p[1] - p[0], p[2] - p[0]
        n = np.cross(e1, e2)
        n /= np.linalg.norm(n)
        t = e1 / np.linalg.norm(e1)
        b = np.cross(n, t)
        loc = np.stack([(p - p[0]) @ t, (p - p[0]) @ b], 1)
        cid, loc = self.new_chart(loc)
        return self.add_geometry(p, np.repeat(n[None], 3, 0), np.repeat(np.append(t, 1)[None], 3, 0),
                                 [0, 1, 2], np.full(3, cid), loc, material)

    def add_emitter_fan(self, center, size, k, material, rng):
        """k triangles around `center` sharing it (a fan of uneven wedges and
        radii, so the triangles' areas differ), horizontal, facing -y, with
        texture uvs over their extent (for a textured emission)."""
        c = np.asarray(center, np.float64)
        cuts = np.sort(rng.uniform(0, 2 * math.pi, k))
        cuts = np.append(cuts, cuts[0] + 2 * math.pi)
        rad = size * rng.uniform(0.4, 1.0, k + 1)
        rad[-1] = rad[0]
        ring = np.stack([c[0] + rad * np.cos(cuts), np.full(k + 1, c[1]),
                         c[2] + rad * np.sin(cuts)], 1)
        pos = np.concatenate([c[None], ring[:-1]])
        tri = []
        for i in range(k):
            # centre, next, this: counter-clockwise seen from below (normal -y)
            tri += [0, 1 + (i + 1) % k, 1 + i]
        n = np.repeat(np.array([[0.0, -1.0, 0.0]]), len(pos), 0)
        tan = np.repeat(np.array([[1.0, 0.0, 0.0, 1.0]]), len(pos), 0)
        lo = pos[:, [0, 2]].min(0)
        ext = np.maximum(pos[:, [0, 2]].max(0) - lo, 1e-6)
        tuv = (pos[:, [0, 2]] - lo) / ext
        cid, loc = self.new_chart(pos[:, [0, 2]])
        return self.add_geometry(pos, n, tan, tri, np.full(len(pos), cid), loc, material, tuv=tuv)

    def add_light(self, **kw):
        self.lights.append(kw)
        return len(self.lights) - 1

    # -------------------------------------------------------------- atlas
    def pack(self, width, height, padding=2):
        """Shelf-packs every chart (tallest first) at one texel density, the
        largest found by bisection that fits, with `padding` texels around
        each chart. When the charts cannot fit as whole texels, places them in
        a grid of sub-texel cells instead."""
        ext = np.array(self.charts, np.float64).reshape(-1, 2)
        order = np.argsort(-ext[:, 1], kind="stable")

        def place(d):
            ws = np.maximum(1, np.ceil(ext[:, 0] * d)).astype(np.int64) + 2 * padding
            hs = np.maximum(1, np.ceil(ext[:, 1] * d)).astype(np.int64) + 2 * padding
            if ws.max() > width:
                return None
            x = y = row_h = 0
            wl, hl = ws[order].tolist(), hs[order].tolist()  # plain ints: a fast loop
            px, py = [0] * len(wl), [0] * len(wl)
            for k in range(len(wl)):
                w, h = wl[k], hl[k]
                if x + w > width:
                    x, y, row_h = 0, y + row_h, 0
                if y + h > height:
                    return None
                px[k], py[k] = x, y
                x += w
                if h > row_h:
                    row_h = h
            pos = np.zeros((len(ext), 2), np.int64)
            pos[order, 0] = px
            pos[order, 1] = py
            return pos

        lo = 0.0
        self.pad_used = padding
        if place(lo) is None:
            # too many charts for whole texels: a grid of equal cells (possibly
            # smaller than a texel), every chart at one density so the largest
            # fits its cell; charts below a texel rasterize through the
            # centroid fallback, as in the reference rasterizer
            n = len(ext)
            gx = int(math.ceil(math.sqrt(n * width / height)))
            gy = int(math.ceil(n / gx))
            cw, chh = width / gx, height / gy
            d = 0.9 * min(cw / max(ext[:, 0].max(), 1e-12), chh / max(ext[:, 1].max(), 1e-12))
            pos = np.zeros((n, 2), np.float64)
            for k, ci in enumerate(order):
                pos[ci] = ((k % gx) * cw + 0.05 * cw, (k // gx) * chh + 0.05 * chh)
            self.pad_used = 0
            self.density = d
            self.chart_pos = pos
            return d
        area = float((ext[:, 0] * ext[:, 1]).sum())
        hi = 2.0 * math.sqrt(width * height / max(area, 1e-12))
        while place(hi) is not None:
            lo, hi = hi, hi * 2
        for _ in range(24):
            mid = 0.5 * (lo + hi)
            if place(mid) is None:
                hi = mid
            else:
                lo = mid
        self.density = lo
        self.chart_pos = place(lo)
        return lo

    def build(self, width, height, padding=2) -> Scene:
        self.pack(width, height, padding)
        s = Scene()
        s.positions = np.concatenate(self.pos).astype(np.float32)
        s.normals = np.concatenate(self.nrm).astype(np.float32)
        s.tangents = np.concatenate(self.tan).astype(np.float32)
        s.texture_uvs = np.concatenate(self.tuv).astype(np.float32)
        loc = np.concatenate(self.local)
        ch = np.concatenate(self.chart_of)
        px = self.chart_pos[ch] + self.pad_used + loc * self.density
        s.lightmap_uvs = np.stack([px[:, 0] / width, px[:, 1] / height], 1).astype(np.float32)
        s.indices = np.concatenate(self.idx).astype(np.uint32)
        s.geometries = self.geoms
        s.materials = self.materials
        s.texels = np.frombuffer(bytes(self.texels), np.uint8).copy()
        s.lights = self.lights
        return s


# ===================================================================== scenes
def box_room_scene(width=256, height=256, padding=2) -> Scene:
    """Box room 4 x 3 x 5 m (x in [-2,2], y in [0,3], z in [-2.5,2.5]), 12
    inward triangles, plus a 0.5 m cube centred at (0, 1.5, 0) with 12 outward
    triangles; one Lambert material of linear albedo 0.7 stored as sRGB8
    (218, i.e. 0.7011 after decode); one point light of intensity 50 at
    (0, 2.5, 0); every face its own chart."""
    b = SceneBuilder()
    g = linear_to_srgb8(0.7)
    m = b.add_material((g, g, g, 255))
    b.add_box((0, 1.5, 0), (2, 1.5, 2.5), material=m, inward=True)
    b.add_box((0, 1.5, 0), (0.25, 0.25, 0.25), material=m)
    b.add_light(type=LIGHT_POINT, position=(0, 2.5, 0), intensity=50.0, color=(1, 1, 1))
    return b.build(width, height, padding)


def random_rotation(rng):
    """Uniform random rotation from three uniforms (Shoemake, Graphics Gems
    III, 1992)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


SYNTH_SEED = 0x5EED0002


def synthetic_room_scene(width=1024, height=1024, seed=SYNTH_SEED, n_pillars=400, n_boxes=2000,
                         n_emissive=64, n_point=4, n_materials=128, segments=32, rings=8,
                         target_tris=262144, padding=2, n_spot=0, emitter_tris=1,
                         textured_emission=False, emitter_y=(13.0, 14.5)) -> Scene:
    """Sponza-scale synthetic room: closed room 40 x 15 x 20 m
    (x in [-20,20], y in [0,15], z in [-10,10]) with `n_pillars` 32-segment
    cylinders (radius U[0.2,0.6], full height, `rings` rings, 30-triangle fan
    caps), `n_boxes` boxes (edges U[0.2,2], uniform rotations), room walls
    subdivided so the total reaches about `target_tris`; `n_materials`
    albedos U[0.2,0.8]^3 (sRGB8), `n_point` point lights (log-U[10,200]),
    one directional light and `n_emissive` emissive triangles (white,
    emission U[1,20]) near the ceiling, each its own area light; `n_spot`
    downward spot lights (log-U[20,400], cones 20-35 / 35-50 degrees).
    emitter_tris > 1 makes each emitter a fan of that many triangles of
    uneven areas (one area light each); textured_emission gives each
    emitter a random 4x4 sRGB texture instead of white; emitters sit at
    heights U[emitter_y]."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    mats = []
    for _ in range(n_materials):
        alb = rng.uniform(0.2, 0.8, 3)
        mats.append(b.add_material(tuple(linear_to_srgb8(a) for a in alb) + (255,)))
    pillar_tris = 2 * segments * rings + 2 * (segments - 2)
    fixed = n_pillars * pillar_tris + 12 * n_boxes + n_emissive
    room_tris = max(12, target_tris - fixed)
    # wall cells so the six walls hold about room_tris triangles
    L = (40.0, 15.0, 20.0)
    walls_area = 2 * (L[0] * L[1] + L[0] * L[2] + L[1] * L[2])
    cell = math.sqrt(walls_area / max(room_tris / 2, 6))
    wall_mat = mats[int(rng.integers(0, n_materials))]
    b.add_box((0, 7.5, 0), (20, 7.5, 10), material=wall_mat, inward=True, cell=cell)
    for _ in range(n_pillars):
        r = rng.uniform(0.2, 0.6)
        cx = rng.uniform(-20 + r, 20 - r)
        cz = rng.uniform(-10 + r, 10 - r)
        b.add_cylinder(cx, cz, r, 0.0, 15.0, segments, rings, mats[int(rng.integers(0, n_materials))])
    centers, halves, Rs, bm = [], [], [], []
    for _ in range(n_boxes):
        halves.append(rng.uniform(0.2, 2.0, 3) / 2)
        centers.append((rng.uniform(-19, 19), rng.uniform(1, 14), rng.uniform(-9, 9)))
        Rs.append(random_rotation(rng))
        bm.append(mats[int(rng.integers(0, n_materials))])
    if n_boxes:
        b.add_boxes(centers, halves, Rs, bm)
    for _ in range(n_emissive):
        if emitter_tris > 1:
            c = np.array([rng.uniform(-19, 19), rng.uniform(*emitter_y), rng.uniform(-9, 9)])
            em_i = float(rng.uniform(1, 20))
            if textured_emission:
                px = rng.integers(60, 256, (4, 4, 3)).astype(np.uint8)
                em = b.add_material(pixels=px, emission=em_i)
            else:
                em = b.add_material((255, 255, 255, 255), emission=em_i)
            g = b.add_emitter_fan(c, rng.uniform(0.3, 1.0), emitter_tris, em, rng)
            b.add_light(type=LIGHT_AREA, geometry_index=g, material_index=em, intensity=em_i)
            continue
        size = rng.uniform(0.3, 1.0)
        c = np.array([rng.uniform(-19, 19), rng.uniform(*emitter_y), rng.uniform(-9, 9)])
        ang = rng.uniform(0, 2 * math.pi)
        pts = [c + size * np.array([math.cos(ang + k * 2 * math.pi / 3), 0,
                                    math.sin(ang + k * 2 * math.pi / 3)]) for k in range(3)]
        pts = [pts[0], pts[1], pts[2]]
        # wind so the face normal points down
        if np.cross(pts[1] - pts[0], pts[2] - pts[0])[1] > 0:
            pts = [pts[0], pts[2], pts[1]]
        em = b.add_material((255, 255, 255, 255), emission=float(rng.uniform(1, 20)))
        g = b.add_triangle(pts, em)
        b.add_light(type=LIGHT_AREA, geometry_index=g, material_index=em,
                    intensity=b.materials[em]["emission"])
    for _ in range(n_point):
        b.add_light(type=LIGHT_POINT,
                    position=(rng.uniform(-18, 18), rng.uniform(1, 14), rng.uniform(-9, 9)),
                    intensity=float(math.exp(rng.uniform(math.log(10), math.log(200)))),
                    color=(1, 1, 1))
    for _ in range(n_spot):
        inner = math.radians(rng.uniform(20, 35))
        outer = inner + math.radians(rng.uniform(5, 15))
        sd = np.array([rng.uniform(-0.3, 0.3), -1.0, rng.uniform(-0.3, 0.3)])
        b.add_light(type=LIGHT_SPOT,
                    position=(rng.uniform(-18, 18), rng.uniform(10, 14), rng.uniform(-9, 9)),
                    direction=tuple(sd / np.linalg.norm(sd)),
                    intensity=float(math.exp(rng.uniform(math.log(20), math.log(400)))),
                    color=(1, 1, 1), cos_inner_cone=math.cos(inner), cos_outer_cone=math.cos(outer))
    d = np.array([rng.uniform(-0.5, 0.5), -1.0, rng.uniform(-0.5, 0.5)])
    b.add_light(type=LIGHT_DIRECTIONAL, direction=tuple(d / np.linalg.norm(d)), intensity=3.0,
                color=(1, 1, 1))
    return b.build(width, height, padding)


def scaled_synthetic_scene(target_tris, width, seed, **kw) -> Scene:
    """The synthetic room scaled to about `target_tris` triangles (pillar and
    box counts proportional to the 262k-triangle configuration)."""
    f = target_tris / 262144.0
    n_p = max(1, int(round(400 * f)))
    n_b = max(4, int(round(2000 * f)))
    n_e = max(2, min(64, int(round(64 * max(f, 0.1)))))
    return synthetic_room_scene(width, width, seed, n_pillars=n_p, n_boxes=n_b, n_emissive=n_e,
                                target_tris=target_tris, **kw)


def stress_scene(width=4096, seed=0x5EED0005) -> Scene:
    """The large stress configuration: the synthetic room with 400 pillars of
    64 segments x 48 rings, 16x the boxes (32000), walls filling up to about
    4M triangles, and 256 lights (184 emissive triangles, 56 point, 15 spot,
    1 directional)."""
    return synthetic_room_scene(width, width, seed, n_pillars=400, n_boxes=32000, n_emissive=184,
                                n_point=56, n_spot=15, segments=64, rings=48,
                                target_tris=4194304)


# -------------------------------------------------- reference test scenes
def quad_scene(with_tangents=True) -> Scene:
    """The +Z quad of baker_test.cpp:16-32 / rasterizer_test.cpp:9-18 with a
    1x1 white RGBA material and a directional light towards -Z."""
    s = Scene()
    s.positions = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    s.normals = np.array([[0, 0, 1]] * 4, np.float32)
    s.texture_uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    s.lightmap_uvs = s.texture_uvs.copy()
    s.tangents = np.array([[1, 0, 0, 1]] * 4, np.float32)
    s.indices = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    flags = GEOM_HAS_NORMALS | GEOM_HAS_TEXTURE_UVS | GEOM_HAS_LIGHTMAP_UVS
    if with_tangents:
        flags |= GEOM_HAS_TANGENTS
    s.geometries = [dict(vertex_offset=0, vertex_count=4, index_offset=0, index_count=6,
                         material_id=0, flags=flags, transform=IDENTITY)]
    s.texels = np.array([255, 255, 255, 255], np.uint8)
    s.materials = [dict(width=1, height=1, channels=4, offset=0, emission=0.0)]
    s.lights = [dict(type=LIGHT_DIRECTIONAL, direction=(0, 0, -1), intensity=1.0, color=(1, 1, 1))]
    return s


def triangles_scene(tris, materials=None, normals=None) -> Scene:
    """Geometries of bare triangles (positions only), as the occlusion tests
    build them (occlusion_test.cpp:20-27, 117-139)."""
    s = Scene()
    pos, idx, geos = [], [], []
    nv = ni = 0
    for gi, g in enumerate(tris):
        if isinstance(g, tuple):
            vtx, ind = np.asarray(g[0], np.float32), np.asarray(g[1], np.uint32)
        else:
            vtx = np.asarray(g, np.float32).reshape(-1, 3)
            ind = np.arange(len(vtx), dtype=np.uint32)
        pos.append(vtx)
        idx.append(ind)
        mid = 0 if materials is None else materials[gi]
        geos.append(dict(vertex_offset=nv, vertex_count=len(vtx), index_offset=ni,
                         index_count=len(ind), material_id=mid, flags=0, transform=IDENTITY))
        nv += len(vtx)
        ni += len(ind)
    s.positions = np.concatenate(pos)
    s.indices = np.concatenate(idx)
    s.geometries = geos
    nmat = 1 if materials is None else max(materials) + 1
    s.materials = [dict(width=0, height=0, channels=0, offset=0, emission=0.0)] * nmat
    return s
