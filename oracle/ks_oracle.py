"""ctypes binding of the CPU oracle (oracle/_build/libks_oracle.so). Test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, as the
checker or as the timed CPU baseline. The scene arrays it consumes are the same flat arrays the GPU
library consumes (keepsake_amd.api.HostScene).
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libks_oracle.so"
# Same sources built with FMA contraction: comparing the two shows how sensitive a reference formula is
# to float evaluation order at a given input.
LIB_PATH_FMA = HERE / "_build" / "libks_oracle_fma.so"


class OraMaterial(C.Structure):
    _fields_ = [
        ("basecolor", C.c_float * 3),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("ior", C.c_float),
        ("specular_r0_mul", C.c_float),
        ("specular_trans", C.c_float),
        ("diffuse_trans", C.c_float),
        ("diffuse_trans_fwd", C.c_float),
        ("emission", C.c_float * 3),
        ("has_emission", C.c_int32),
    ]


class OraLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", C.c_float * 3), ("vec", C.c_float * 3)]


class OraSceneDesc(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p),
        ("n_vertices", C.c_uint32),
        ("normals", C.c_void_p),
        ("indices", C.c_void_p),
        ("n_tris", C.c_uint32),
        ("tri_geom", C.c_void_p),
        ("geom_material", C.c_void_p),
        ("n_geoms", C.c_uint32),
        ("materials", C.c_void_p),
        ("n_materials", C.c_uint32),
        ("lights", C.c_void_p),
        ("n_lights", C.c_uint32),
        ("geom_flags", C.c_void_p),
        ("sky_map", C.c_void_p),
        ("texcoords", C.c_void_p),
        ("textures", C.c_void_p),
        ("n_textures", C.c_uint32),
        ("material_fields", C.c_void_p),
    ]


class OraTexture(C.Structure):
    _fields_ = [("data", C.c_void_p)] + [(n, C.c_int32) for n in ("width", "height", "channels", "data_type",
                                                                 "color_space", "levels", "build_mipmaps")]


class OraField(C.Structure):
    _fields_ = [
        ("texture", C.c_int32),
        ("sampler", C.c_int32),
        ("cubic_kernel", C.c_int32),
        ("anisotropy", C.c_float),
        ("wrap_u", C.c_int32),
        ("wrap_v", C.c_int32),
        ("flip_v", C.c_int32),
        ("swizzle", C.c_int32 * 4),
        ("uv_scale", C.c_float * 2),
        ("uv_offset", C.c_float * 2),
        ("scale", C.c_float * 4),
        ("value", C.c_float * 4),
    ]


ORA_MATERIAL_FIELDS_SIZE = 8 * C.sizeof(OraField) + 4 + C.sizeof(OraField) + 4 + C.sizeof(OraField) + 12 + 36 + 4 + C.sizeof(OraField)


class OraSkyMap(C.Structure):
    _fields_ = [
        ("rgb", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("to_world", C.c_float * 9),
        ("transform_y_up", C.c_int32),
        ("strength", C.c_float),
    ]


class OraCamera(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3),
        ("target", C.c_float * 3),
        ("up", C.c_float * 3),
        ("vfov_deg", C.c_float),
        ("aspect", C.c_float),
    ]


class OraParams(C.Structure):
    _fields_ = [
        ("bounces", C.c_int32),
        ("render_width", C.c_int32),
        ("render_height", C.c_int32),
        ("crop_start_x", C.c_int32),
        ("crop_start_y", C.c_int32),
        ("crop_width", C.c_int32),
        ("crop_height", C.c_int32),
        ("spp", C.c_int32),
        ("rng_seed", C.c_int32),
        ("include_background", C.c_int32),
        ("backdrop", C.c_float * 3),
        ("clamp_direct", C.c_float),
        ("clamp_indirect", C.c_float),
        ("mis_mode", C.c_int32),
        ("scale_ray_diff", C.c_int32),
        ("pixel_filter", C.c_int32),
        ("filter_width", C.c_float),
        ("filter_sigma", C.c_float),
    ]


_P = C.c_void_p
_SIGS = {
    "ora_scene_create": ([C.POINTER(OraSceneDesc)], _P),
    "ora_scene_free": ([_P], None),
    "ora_intersect": ([_P, _P, C.c_int64, _P, C.c_int], C.c_int),
    "ora_intersect_brute": ([_P, _P, C.c_int64, _P, C.c_int], C.c_int),
    "ora_occluded": ([_P, _P, C.c_int64, _P, C.c_int], C.c_int),
    "ora_light_table": ([_P, C.POINTER(C.c_int32), _P, _P, _P], C.c_int),
    "ora_sky_eval": ([_P, _P, C.c_int64, _P, _P], C.c_int),
    "ora_sky_sample": ([_P, _P, C.c_int64, _P, _P, _P], C.c_int),
    "ora_field_eval": ([_P, _P, _P, C.c_int64, _P], C.c_int),
    "ora_tex_tables": ([_P, _P], None),
    "ora_primary_surface": ([_P, C.POINTER(OraCamera), C.POINTER(OraParams), _P, _P, _P, C.c_int64, _P], C.c_int),
    "ora_texture_level": ([_P, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _P], C.c_int),
    "ora_render": ([_P, C.POINTER(OraCamera), C.POINTER(OraParams), _P, _P, C.c_int], C.c_int),
    "ora_trace_samples": ([_P, C.POINTER(OraCamera), C.POINTER(OraParams), _P, _P, _P, C.c_int64, _P, C.c_int], C.c_int),
    "ora_sobol": ([C.c_int] * 8 + [_P, _P, _P], C.c_int),
    "ora_murmur64a": ([_P, C.c_int, C.c_uint64], C.c_uint64),
    "ora_mix_bits": ([C.c_uint64], C.c_uint64),
    "ora_owen_scramble": ([C.c_uint32, C.c_uint32], C.c_uint32),
    "ora_sobol_matrices": ([_P], None),
    "ora_camera_rays": ([C.POINTER(OraCamera), C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_int64, _P], C.c_int),
    "ora_camera_matrices": ([C.POINTER(OraCamera), _P, _P], C.c_int),
    "ora_bsdf_eval": ([C.POINTER(OraMaterial), _P, _P, C.c_int64, _P, _P], C.c_int),
    "ora_bsdf_sample": ([C.POINTER(OraMaterial), _P, _P, C.c_int64, _P, _P, _P], C.c_int),
    "ora_ggx_eval": ([C.c_float, C.c_float, C.c_int, _P, _P, C.c_int64, _P], C.c_int),
    "ora_ggx_sample": ([C.c_float, C.c_float, _P, _P, C.c_int64, _P], C.c_int),
    "ora_fresnel": ([C.c_int, _P, C.c_float, C.c_int64, _P], C.c_int),
    "ora_bsdf_lobes": ([C.POINTER(OraMaterial), _P, C.c_int64, _P], C.c_int),
    "ora_bsdf_lobe": ([C.POINTER(OraMaterial), C.c_int, _P, _P, C.c_int64, _P], C.c_int),
    "ora_light_sample": ([_P, C.c_int, _P, _P, C.c_int64, _P, _P, _P, _P], C.c_int),
    "ora_light_pdf": ([_P, C.c_int, _P, _P, C.c_int64, _P], C.c_int),
    "ora_light_select": ([_P, _P, C.c_int64, _P, _P], C.c_int),
    "ora_pixel_filter": ([C.POINTER(OraParams), _P, _P, C.c_int64, _P, _P], C.c_int),
}

_libs: dict = {}


def load(variant: str = "") -> C.CDLL:
    """variant "" = no FMA contraction (default), "fma" = FMA-contracted build of the same sources."""
    if variant not in _libs:
        path = LIB_PATH_FMA if variant == "fma" else LIB_PATH
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        lib = C.CDLL(str(path))
        for n, (a, r) in _SIGS.items():
            f = getattr(lib, n)
            f.argtypes = a
            f.restype = r
        _libs[variant] = lib
    return _libs[variant]


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def threads_default() -> int:
    return int(os.environ.get("KS_ORACLE_THREADS", "0"))


def camera(cam: dict) -> OraCamera:
    c = OraCamera()
    for i in range(3):
        c.pos[i] = float(cam["pos"][i])
        c.target[i] = float(cam["target"][i])
        c.up[i] = float(cam.get("up", (0.0, 1.0, 0.0))[i])
    c.vfov_deg = float(cam["vfov"])
    c.aspect = float(cam["aspect"])
    return c


def params(**kw) -> OraParams:
    p = OraParams()
    p.bounces = kw.get("bounces", 1)
    p.render_width = kw.get("render_width", 256)
    p.render_height = kw.get("render_height", 256)
    p.crop_start_x = kw.get("crop_start_x", 0)
    p.crop_start_y = kw.get("crop_start_y", 0)
    p.crop_width = kw.get("crop_width", 0)
    p.crop_height = kw.get("crop_height", 0)
    p.spp = kw.get("spp", 1)
    p.rng_seed = kw.get("rng_seed", 0)
    p.include_background = int(kw.get("include_background", True))
    bd = kw.get("backdrop", (0.0, 0.0, 0.0))
    for i in range(3):
        p.backdrop[i] = float(bd[i])
    p.clamp_direct = kw.get("clamp_direct", 0.0)
    p.clamp_indirect = kw.get("clamp_indirect", 10.0)
    p.mis_mode = kw.get("mis_mode", 0)
    p.scale_ray_diff = int(kw.get("scale_ray_diff", True))
    p.pixel_filter = kw.get("pixel_filter", 0)
    p.filter_width = kw.get("filter_width", 0.0)
    p.filter_sigma = kw.get("filter_sigma", 0.0)
    return p


def material(rec) -> OraMaterial:
    m = OraMaterial()
    b = np.asarray(rec).tobytes()
    assert len(b) == C.sizeof(m)
    C.memmove(C.byref(m), b, C.sizeof(m))
    return m


class OracleScene:
    def __init__(self, vertices, indices, tri_geom, geom_material, materials, lights, normals=None, geom_flags=None,
                 sky_map=None, texcoords=None, textures=None, material_fields=None):
        """sky_map: dict(rgb=(H, W, 3), to_world=3x3, y_up=bool, strength=float) for a type-3 light.
        texcoords: (V, 2); textures: dicts with "data" (an (H, W, C) uint8 / uint16 / float32 array, or a
        list of them, one per level), "srgb", "mipmaps"; material_fields: material index -> a ctypes
        structure with the ora_material_fields layout."""
        lib = load()
        self.lib = lib
        self._arrs = [
            np.ascontiguousarray(vertices, np.float32),
            np.ascontiguousarray(indices, np.uint32),
            np.ascontiguousarray(tri_geom, np.uint32),
            np.ascontiguousarray(geom_material, np.uint32),
            np.ascontiguousarray(materials),
            np.ascontiguousarray(lights),
            None if normals is None else np.ascontiguousarray(normals, np.float32),
            None if geom_flags is None else np.ascontiguousarray(geom_flags, np.uint8),
        ]
        v, idx, tg, gm, mats, lts, nrm, gf = self._arrs
        d = OraSceneDesc()
        d.vertices = v.ctypes.data
        d.n_vertices = v.shape[0]
        d.normals = nrm.ctypes.data if nrm is not None else None
        d.indices = idx.ctypes.data
        d.n_tris = idx.shape[0]
        d.tri_geom = tg.ctypes.data
        d.geom_material = gm.ctypes.data
        d.n_geoms = gm.shape[0]
        d.materials = mats.ctypes.data
        d.n_materials = mats.shape[0]
        d.lights = lts.ctypes.data if lts.shape[0] else None
        d.n_lights = lts.shape[0]
        d.geom_flags = gf.ctypes.data if gf is not None else None
        if sky_map is not None:
            rgb = np.ascontiguousarray(sky_map["rgb"], np.float32)
            sm = OraSkyMap()
            sm.rgb = rgb.ctypes.data
            sm.height, sm.width = rgb.shape[0], rgb.shape[1]
            m = np.asarray(sky_map.get("to_world", np.eye(3)), np.float32).reshape(9)
            for k in range(9):
                sm.to_world[k] = float(m[k])
            sm.transform_y_up = 1 if sky_map.get("y_up", True) else 0
            sm.strength = float(sky_map.get("strength", 1.0))
            self._arrs += [rgb, sm]
            d.sky_map = C.addressof(sm)
        if texcoords is not None:
            tc = np.ascontiguousarray(texcoords, np.float32)
            self._arrs.append(tc)
            d.texcoords = tc.ctypes.data
        if textures:
            types = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}
            arr = (OraTexture * len(textures))()
            for i, t in enumerate(textures):
                levels = [np.asarray(x) for x in (t["data"] if isinstance(t["data"], (list, tuple)) else [t["data"]])]
                lv0 = levels[0] if levels[0].ndim == 3 else levels[0][..., None]
                blob = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(x).reshape(-1) for x in levels]))
                self._arrs.append(blob)
                arr[i].data = blob.ctypes.data
                arr[i].height, arr[i].width, arr[i].channels = lv0.shape
                arr[i].data_type = types[lv0.dtype]
                arr[i].color_space = 1 if t.get("srgb", False) else 0
                arr[i].levels = len(levels)
                arr[i].build_mipmaps = 1 if t.get("mipmaps", True) else 0
            self._arrs.append(arr)
            d.textures = C.addressof(arr)
            d.n_textures = len(textures)
        if material_fields:
            ptrs = (C.c_void_p * mats.shape[0])()
            for m, mf in material_fields.items():
                assert C.sizeof(mf) == ORA_MATERIAL_FIELDS_SIZE
                buf = C.create_string_buffer(C.string_at(C.addressof(mf), C.sizeof(mf)), C.sizeof(mf))
                self._arrs.append(buf)
                ptrs[m] = C.addressof(buf)
            self._arrs.append(ptrs)
            d.material_fields = C.addressof(ptrs)
        self.h = lib.ora_scene_create(C.byref(d))
        if not self.h:
            raise ValueError("ora_scene_create rejected the scene")

    @classmethod
    def from_host_scene(cls, hs) -> "OracleScene":
        return cls(hs.vertices, hs.indices, hs.tri_geom, hs.geom_material, hs.materials, hs.lights, hs.normals,
                   getattr(hs, "geom_flags", None), getattr(hs, "sky_map", None), getattr(hs, "texcoords", None),
                   getattr(hs, "textures", None), getattr(hs, "material_fields", None))

    def close(self):
        if self.h:
            self.lib.ora_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def intersect(self, rays, brute=False, threads=None):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((rays.shape[0], 4), np.float32)
        fn = self.lib.ora_intersect_brute if brute else self.lib.ora_intersect
        fn(self.h, _p(rays), rays.shape[0], _p(out), threads_default() if threads is None else threads)
        return out

    def occluded(self, rays, threads=None):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(rays.shape[0], np.uint8)
        self.lib.ora_occluded(self.h, _p(rays), rays.shape[0], _p(out), threads_default() if threads is None else threads)
        return out

    def light_table(self):
        n = C.c_int32(0)
        self.lib.ora_light_table(self.h, C.byref(n), None, None, None)
        prim = np.zeros(n.value, np.int32)
        power = np.zeros(n.value, np.float32)
        cdf = np.zeros(n.value + 1, np.float32)
        self.lib.ora_light_table(self.h, C.byref(n), _p(prim), _p(power), _p(cdf))
        return prim, power, cdf

    def light_sample(self, light, p, u):
        """Light `light` of the flat table sampled at points p (n, 3) with u (n, 2): wi, shadow-ray
        distance, solid-angle pdf, Le / pdf."""
        p = np.ascontiguousarray(np.broadcast_to(p, (len(u), 3)), np.float32)
        u = np.ascontiguousarray(u, np.float32)
        n = u.shape[0]
        wi, d, pdf, b = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        assert self.lib.ora_light_sample(self.h, light, _p(p), _p(u), n, _p(wi), _p(d), _p(pdf), _p(b)) == 0
        return wi, d, pdf, b

    def light_pdf(self, light, wi, dist):
        wi = np.ascontiguousarray(wi, np.float32)
        dist = np.ascontiguousarray(dist, np.float32)
        out = np.zeros(wi.shape[0], np.float32)
        assert self.lib.ora_light_pdf(self.h, light, _p(wi), _p(dist), wi.shape[0], _p(out)) == 0
        return out

    def light_select(self, u):
        u = np.ascontiguousarray(u, np.float32)
        idx = np.zeros(u.shape[0], np.int32)
        prob = np.zeros(u.shape[0], np.float32)
        assert self.lib.ora_light_select(self.h, _p(u), u.shape[0], _p(idx), _p(prob)) == 0
        return idx, prob

    def sky_eval(self, dirs):
        """The sky-map light's radiance (n, 3) and solid-angle pdf (n) at world directions (n, 3)."""
        dirs = np.ascontiguousarray(dirs, np.float32)
        le = np.zeros((dirs.shape[0], 3), np.float32)
        pdf = np.zeros(dirs.shape[0], np.float32)
        assert self.lib.ora_sky_eval(self.h, _p(dirs), dirs.shape[0], _p(le), _p(pdf)) == 0
        return le, pdf

    def sky_sample(self, u):
        """Sky-map light samples from u (n, 2): directions (n, 3), Le / pdf (n, 3), pdf (n)."""
        u = np.ascontiguousarray(u, np.float32)
        wi = np.zeros((u.shape[0], 3), np.float32)
        b = np.zeros((u.shape[0], 3), np.float32)
        pdf = np.zeros(u.shape[0], np.float32)
        assert self.lib.ora_sky_sample(self.h, _p(u), u.shape[0], _p(wi), _p(b), _p(pdf)) == 0
        return wi, b, pdf

    def field_eval(self, field, uvd):
        """ShaderField value (n, 4) of `field` (a ctypes structure with the ora_field layout) at rows
        (u, v, dudx, dvdx, dudy, dvdy)."""
        uvd = np.ascontiguousarray(uvd, np.float32)
        out = np.zeros((uvd.shape[0], 4), np.float32)
        f = OraField.from_buffer_copy(C.string_at(C.addressof(field), C.sizeof(OraField)))
        assert self.lib.ora_field_eval(self.h, C.byref(f), _p(uvd), uvd.shape[0], _p(out)) == 0
        return out

    def primary_surface(self, cam: dict, prm: OraParams, px, py, sample):
        """(u, v, dudx, dvdx, dudy, dvdy, shading n, shading t, geometric n, prim bits) at the first hit."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        out = np.zeros((px.shape[0], 16), np.float32)
        c = camera(cam)
        assert self.lib.ora_primary_surface(self.h, C.byref(c), C.byref(prm), _p(px), _p(py), _p(sample),
                                            px.shape[0], _p(out)) == 0
        return out

    def texture_level(self, tex, level):
        w, h = C.c_int(0), C.c_int(0)
        assert self.lib.ora_texture_level(self.h, tex, level, C.byref(w), C.byref(h), None) == 0
        nc = self.lib.ora_texture_level(self.h, tex, -1, C.byref(w), C.byref(h), None)
        self.lib.ora_texture_level(self.h, tex, level, C.byref(w), C.byref(h), None)
        out = np.zeros((h.value, w.value, nc), np.float32)
        assert self.lib.ora_texture_level(self.h, tex, level, C.byref(w), C.byref(h), _p(out)) == 0
        return out

    def render(self, cam: dict, prm: OraParams, threads=None):
        W, H = prm.render_width, prm.render_height
        rgb = np.zeros((H, W, 3), np.float64)
        w = np.zeros((H, W, 2), np.float64)
        c = camera(cam)
        self.lib.ora_render(self.h, C.byref(c), C.byref(prm), _p(rgb), _p(w), threads_default() if threads is None else threads)
        return rgb, w

    def trace_samples(self, cam: dict, prm: OraParams, px, py, sample, threads=None):
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        out = np.zeros((px.shape[0], 4), np.float32)
        c = camera(cam)
        self.lib.ora_trace_samples(
            self.h, C.byref(c), C.byref(prm), _p(px), _p(py), _p(sample), px.shape[0], _p(out),
            threads_default() if threads is None else threads,
        )
        return out


def sobol(width, height, spp, seed, px, py, sample, kinds):
    lib = load()
    kinds = np.ascontiguousarray(kinds, np.int32)
    nv = int(kinds.sum())
    ou = np.zeros(nv, np.uint32)
    of = np.zeros(nv, np.float32)
    lib.ora_sobol(width, height, spp, seed, px, py, sample, kinds.shape[0], _p(kinds), _p(ou), _p(of))
    return ou, of


def murmur64a(data: bytes, seed: int = 0) -> int:
    lib = load()
    buf = C.create_string_buffer(data, len(data))
    return lib.ora_murmur64a(C.cast(buf, C.c_void_p), len(data), seed)


def mix_bits(v: int) -> int:
    return load().ora_mix_bits(v)


def owen_scramble(v: int, seed: int) -> int:
    return load().ora_owen_scramble(v, seed)


def sobol_matrices() -> np.ndarray:
    out = np.zeros(104, np.uint32)
    load().ora_sobol_matrices(_p(out))
    return out


def camera_rays(cam, width, height, spp, seed, px, py, sample):
    lib = load()
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    sample = np.ascontiguousarray(sample, np.int32)
    out = np.zeros((px.shape[0], 6), np.float32)
    c = camera(cam)
    lib.ora_camera_rays(C.byref(c), width, height, spp, seed, _p(px), _p(py), _p(sample), px.shape[0], _p(out))
    return out


def camera_matrices(cam):
    p2w = np.zeros(16, np.float32)
    pos = np.zeros(3, np.float32)
    c = camera(cam)
    load().ora_camera_matrices(C.byref(c), _p(p2w), _p(pos))
    return p2w.reshape(4, 4), pos


def tex_tables():
    srgb = np.zeros(256, np.float32)
    ewa = np.zeros(128, np.float32)
    load().ora_tex_tables(_p(srgb), _p(ewa))
    return srgb, ewa


def bsdf_eval(mat, wo, wi, variant: str = ""):
    m = material(mat)
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    f = np.zeros((wo.shape[0], 3), np.float32)
    pdf = np.zeros(wo.shape[0], np.float32)
    load(variant).ora_bsdf_eval(C.byref(m), _p(wo), _p(wi), wo.shape[0], _p(f), _p(pdf))
    return f, pdf


def bsdf_sample(mat, wo, u, variant: str = ""):
    m = material(mat)
    wo = np.ascontiguousarray(wo, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    n = wo.shape[0]
    wi = np.zeros((n, 3), np.float32)
    beta = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    load(variant).ora_bsdf_sample(C.byref(m), _p(wo), _p(u), n, _p(wi), _p(beta), _p(pdf))
    return wi, beta, pdf


GGX_WHAT = {"D": 0, "lambda": 1, "G1": 2, "G2": 3}


def ggx(what, ax, ay, a, b=None, variant: str = ""):
    """GGX building block of the oracle (microfacet.h:20-153): D(a), lambda(a), G1(a) or G2(a, b)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(a if b is None else b, np.float32)
    out = np.zeros(a.shape[0], np.float32)
    assert load(variant).ora_ggx_eval(ax, ay, GGX_WHAT[what], _p(a), _p(b), a.shape[0], _p(out)) == 0
    return out


def ggx_sample(ax, ay, wo, u):
    wo = np.ascontiguousarray(np.broadcast_to(wo, (len(u), 3)), np.float32)
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((u.shape[0], 3), np.float32)
    assert load().ora_ggx_sample(ax, ay, _p(wo), _p(u), u.shape[0], _p(out)) == 0
    return out


def fresnel_dielectric(c, eta):
    c = np.ascontiguousarray(c, np.float32)
    out = np.zeros(c.shape[0], np.float32)
    assert load().ora_fresnel(0, _p(c), eta, c.shape[0], _p(out)) == 0
    return out


def fresnel_schlick(c):
    c = np.ascontiguousarray(c, np.float32)
    out = np.zeros(c.shape[0], np.float32)
    assert load().ora_fresnel(1, _p(c), 0.0, c.shape[0], _p(out)) == 0
    return out


def hg_phase(c, g):
    c = np.ascontiguousarray(c, np.float32)
    out = np.zeros(c.shape[0], np.float32)
    assert load().ora_fresnel(2, _p(c), g, c.shape[0], _p(out)) == 0
    return out


def bsdf_lobe_weights(mat, wo):
    m = material(mat)
    wo = np.ascontiguousarray(wo, np.float32)
    out = np.zeros((wo.shape[0], 4), np.float32)
    assert load().ora_bsdf_lobes(C.byref(m), _p(wo), wo.shape[0], _p(out)) == 0
    return out


def bsdf_lobe(mat, lobe, wo, wi):
    """One lobe's f*cos (0 diffuse, 1 diffuse transmission, 2 metallic, 3 dielectric); lobe -1 gives the
    mixture pdf without eval's zero-f gating (first column)."""
    m = material(mat)
    wo = np.ascontiguousarray(np.broadcast_to(wo, np.shape(wi)), np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    out = np.zeros((wi.shape[0], 3), np.float32)
    assert load().ora_bsdf_lobe(C.byref(m), lobe, _p(wo), _p(wi), wi.shape[0], _p(out)) == 0
    return out


def pixel_filter(prm: OraParams, u=None, xy=None):
    """The params' pixel filter: (offsets sampled from u, values at xy, sample-table size)."""
    n = len(u) if u is not None else len(xy)
    u = None if u is None else np.ascontiguousarray(u, np.float32)
    xy = None if xy is None else np.ascontiguousarray(xy, np.float32)
    off = np.zeros((n, 2), np.float32)
    val = np.zeros(n, np.float32)
    tn = load().ora_pixel_filter(C.byref(prm), None if u is None else _p(u), None if xy is None else _p(xy), n,
                                 _p(off), _p(val))
    return off, val, tn
