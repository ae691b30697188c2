"""ctypes bindings of libkeepsake_amd (include/keepsake_amd.h).

Python-side mirror of the reference's small_pt entry points for tests and benchmarks. Every call goes
through the C ABI of the gfx950 library; there is no CPU fallback: loading fails loudly when the
library is missing, and opening a device fails loudly when no GPU is present.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libkeepsake_amd.so"

KS_OK = 0
KS_ERR_INVALID = -1
KS_ERR_HIP = -2
KS_ERR_OOM = -3
KS_ERR_UNSUPPORTED = -4
KS_ERR_NO_DEVICE = -5

KS_FILTER_BOX = 0
KS_FILTER_GAUSSIAN = 1
KS_FILTER_BLACKMAN_HARRIS = 2

KS_LIGHT_SKY_AMBIENT = 0
KS_LIGHT_DIRECTIONAL = 1
KS_LIGHT_POINT = 2
KS_LIGHT_SKY_MAP = 3
KS_MIS_REFERENCE = 0
KS_MIS_CONSISTENT = 1
KS_GUIDE_OFF = 0
KS_GUIDE_TRAIN_AND_SAMPLE = 1

KERNEL_CLASSES = ("raygen", "closest", "shade", "shadow", "film", "guide", "tail")


class KsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"keepsake_amd error {status}: {message}")
        self.status = status


class MaterialDesc(C.Structure):
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


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", C.c_float * 3), ("vec", C.c_float * 3)]


class SkyMapDesc(C.Structure):
    _fields_ = [
        ("rgb", C.POINTER(C.c_float)),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("to_world", C.c_float * 9),
        ("transform_y_up", C.c_int32),
        ("strength", C.c_float),
    ]


KS_TEX_U8, KS_TEX_U16, KS_TEX_F32 = 0, 1, 2
KS_COLORSPACE_LINEAR, KS_COLORSPACE_SRGB = 0, 1
KS_SAMPLER_NEAREST, KS_SAMPLER_LINEAR, KS_SAMPLER_CUBIC, KS_SAMPLER_EWA = 0, 1, 2, 3
KS_WRAP_REPEAT, KS_WRAP_CLAMP = 0, 1
KS_CUBIC_MITCHELL, KS_CUBIC_BSPLINE, KS_CUBIC_CATMULL_ROM = 0, 1, 2
KS_NORMAL_MAP_TANGENT, KS_NORMAL_MAP_WORLD = 0, 1


class TextureDesc(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("channels", C.c_int32),
        ("data_type", C.c_int32),
        ("color_space", C.c_int32),
        ("levels", C.c_int32),
        ("build_mipmaps", C.c_int32),
    ]


class FieldDesc(C.Structure):
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


class MaterialFields(C.Structure):
    _fields_ = [
        ("basecolor", FieldDesc),
        ("roughness", FieldDesc),
        ("metallic", FieldDesc),
        ("ior", FieldDesc),
        ("specular_r0_mul", FieldDesc),
        ("specular_trans", FieldDesc),
        ("diffuse_trans", FieldDesc),
        ("diffuse_trans_fwd", FieldDesc),
        ("has_emission", C.c_int32),
        ("emission", FieldDesc),
        ("has_normal_map", C.c_int32),
        ("normal_map", FieldDesc),
        ("normal_space", C.c_int32),
        ("normal_range01", C.c_int32),
        ("normal_strength", C.c_float),
        ("normal_to_world", C.c_float * 9),
        ("has_opacity_map", C.c_int32),
        ("opacity", FieldDesc),
    ]


_TEX_DTYPES = {np.dtype(np.uint8): KS_TEX_U8, np.dtype(np.uint16): KS_TEX_U16, np.dtype(np.float32): KS_TEX_F32}


def field_desc(value=None, texture=-1, sampler=KS_SAMPLER_LINEAR, cubic_kernel=KS_CUBIC_MITCHELL, anisotropy=8.0,
               wrap_u=KS_WRAP_REPEAT, wrap_v=KS_WRAP_REPEAT, flip_v=True, swizzle=None, uv_scale=(1.0, 1.0),
               uv_offset=(0.0, 0.0), scale=(1.0, 1.0, 1.0, 1.0)) -> FieldDesc:
    """A ShaderField: the constant `value` (texture = -1) or a texture lookup (TextureField defaults:
    flip_v on, identity swizzle, unit scale)."""
    f = FieldDesc()
    f.texture = texture
    f.sampler = sampler
    f.cubic_kernel = cubic_kernel
    f.anisotropy = anisotropy
    f.wrap_u, f.wrap_v = wrap_u, wrap_v
    f.flip_v = 1 if flip_v else 0
    sw = list(swizzle) if swizzle is not None else [0, 1, 2, 3]
    for k in range(4):
        f.swizzle[k] = sw[k] if k < len(sw) else -1
    f.uv_scale[0], f.uv_scale[1] = uv_scale
    f.uv_offset[0], f.uv_offset[1] = uv_offset
    sc = list(np.broadcast_to(np.asarray(scale, np.float32), (4,))) if np.ndim(scale) == 0 else list(scale) + [1.0] * (4 - len(scale))
    for k in range(4):
        f.scale[k] = float(sc[k])
    if value is not None:
        v = np.atleast_1d(np.asarray(value, np.float32))
        for k in range(min(4, v.size)):
            f.value[k] = float(v[k])
    return f


def material_fields(base: np.void | None = None, **fields) -> MaterialFields:
    """Textured material: every closure input a field. Fields not given take the constant of `base` (a
    MATERIAL_DTYPE record); keyword arguments are FieldDesc values or dicts for field_desc().
    normal_map / opacity (FieldDesc) turn the maps on; normal_space, normal_range01, normal_strength,
    normal_to_world set the normal map's options."""
    mf = MaterialFields()
    names = ["basecolor", "roughness", "metallic", "ior", "specular_r0_mul", "specular_trans", "diffuse_trans",
             "diffuse_trans_fwd"]
    for n in names + ["emission"]:
        f = fields.get(n)
        if f is None:
            val = base[n] if base is not None else (0.0 if n != "ior" else 1.5)
            f = field_desc(value=val)
        elif isinstance(f, dict):
            f = field_desc(**f)
        setattr(mf, n, f)
    mf.has_emission = 1 if (fields.get("emission") is not None or (base is not None and base["has_emission"])) else 0
    if fields.get("normal_map") is not None:
        f = fields["normal_map"]
        mf.normal_map = field_desc(**f) if isinstance(f, dict) else f
        mf.has_normal_map = 1
        mf.normal_space = fields.get("normal_space", KS_NORMAL_MAP_TANGENT)
        mf.normal_range01 = 1 if fields.get("normal_range01", True) else 0
        mf.normal_strength = fields.get("normal_strength", 1.0)
        m = np.asarray(fields.get("normal_to_world", np.eye(3)), np.float32).reshape(9)
        for k in range(9):
            mf.normal_to_world[k] = float(m[k])
    if fields.get("opacity") is not None:
        f = fields["opacity"]
        mf.opacity = field_desc(**f) if isinstance(f, dict) else f
        mf.has_opacity_map = 1
    return mf


class SceneDesc(C.Structure):
    _fields_ = [
        ("vertices", C.POINTER(C.c_float)),
        ("n_vertices", C.c_uint32),
        ("normals", C.POINTER(C.c_float)),
        ("indices", C.POINTER(C.c_uint32)),
        ("n_tris", C.c_uint32),
        ("tri_geom", C.POINTER(C.c_uint32)),
        ("geom_material", C.POINTER(C.c_uint32)),
        ("n_geoms", C.c_uint32),
        ("materials", C.POINTER(MaterialDesc)),
        ("n_materials", C.c_uint32),
        ("lights", C.POINTER(LightDesc)),
        ("n_lights", C.c_uint32),
        ("geom_flags", C.POINTER(C.c_uint8)),
        ("sky_map", C.POINTER(SkyMapDesc)),
        ("texcoords", C.POINTER(C.c_float)),
        ("textures", C.POINTER(TextureDesc)),
        ("n_textures", C.c_uint32),
        ("material_fields", C.POINTER(C.POINTER(MaterialFields))),
    ]


KS_GEOM_ONE_SIDED = 1
KS_GEOM_FLAT = 2
KS_GEOM_NO_TEXCOORD = 4


class BvhOptions(C.Structure):
    _fields_ = [
        ("max_prims_in_node", C.c_uint32),
        ("bucket_count", C.c_uint32),
        ("trav_weight", C.c_float),
        ("alpha", C.c_float),
        ("threads", C.c_int32),
        ("branching", C.c_int32),
        ("builder", C.c_int32),
    ]


KS_BVH_BUILD_HOST_SBVH = 0
KS_BVH_BUILD_DEVICE = 1


class CameraDesc(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3),
        ("target", C.c_float * 3),
        ("up", C.c_float * 3),
        ("vfov_deg", C.c_float),
        ("aspect", C.c_float),
        ("use_matrix", C.c_int32),
        ("to_world", C.c_float * 16),
    ]


class SmallPTParams(C.Structure):
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
        ("scale_ray_diff", C.c_int32),
        ("clamp_direct", C.c_float),
        ("clamp_indirect", C.c_float),
        ("spp_prog_interval", C.c_int32),
        ("mis_mode", C.c_int32),
        ("guide_mode", C.c_int32),
        ("queue_size", C.c_int32),
        ("pixel_filter", C.c_int32),
        ("filter_width", C.c_float),
        ("filter_sigma", C.c_float),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("wall_ms", C.c_double),
        ("kernel_ms", C.c_double * 7),
        ("launches", C.c_uint64 * 7),
        ("paths", C.c_uint64),
        ("rays_closest", C.c_uint64),
        ("rays_shadow", C.c_uint64),
        ("rays_tail", C.c_uint64),
        ("node_visits_closest", C.c_uint64),
        ("tri_tests_closest", C.c_uint64),
        ("node_visits_shadow", C.c_uint64),
        ("tri_tests_shadow", C.c_uint64),
        ("nonfinite", C.c_uint64),
        ("stack_overflows", C.c_uint64),
        ("iterations", C.c_int32),
        ("queue_size", C.c_int32),
        ("guide_samples", C.c_uint64),
        ("trace_blocks_per_cu", C.c_int32),
        ("batch_chunks", C.c_int32),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("kernel_ms", "launches")}
        d["kernel_ms"] = {n: self.kernel_ms[i] for i, n in enumerate(KERNEL_CLASSES)}
        d["launches"] = {n: int(self.launches[i]) for i, n in enumerate(KERNEL_CLASSES)}
        return d


class SceneInfo(C.Structure):
    _fields_ = [
        ("n_tris", C.c_uint32),
        ("n_vertices", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("n_refs", C.c_uint32),
        ("n_lights", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("device_bytes", C.c_uint64),
        ("build_ms", C.c_double),
        ("bound_lo", C.c_float * 3),
        ("bound_hi", C.c_float * 3),
        ("branching", C.c_uint32),
        ("builder", C.c_int32),
        ("bvh_build_ms", C.c_double),
    ]


class GuideConfig(C.Structure):
    """ks_guide_config; zero fields take the library defaults."""

    _fields_ = [
        ("log2_table_size", C.c_int32),
        ("base_resolution", C.c_float),
        ("max_resolution", C.c_float),
        ("learning_rate", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("guide_fraction", C.c_float),
        ("train_batch", C.c_int32),
        ("max_steps_per_pass", C.c_int32),
        ("seed", C.c_uint64),
        ("train_spp", C.c_int32),
        ("weight_clamp", C.c_float),
        ("inference_fp8", C.c_int32),
        ("max_guided_bounce", C.c_int32),
    ]


class GuideStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("samples", C.c_int64),
        ("last_loss", C.c_double),
        ("last_weight", C.c_double),
        ("last_loss_cosine", C.c_double),
    ]


class BatchView(C.Structure):
    _fields_ = [("camera", CameraDesc), ("params", SmallPTParams), ("film", C.c_void_p)]


class BatchOptions(C.Structure):
    _fields_ = [("queue_size", C.c_int32), ("trace_blocks_per_cu", C.c_int32), ("disable_graph", C.c_int32),
                ("max_chunk_items", C.c_int32)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_void_p)

_P = C.c_void_p
_SIGS = {
    "ks_last_error": ([], C.c_char_p),
    "ks_device_count": ([C.POINTER(C.c_int32)], C.c_int),
    "ks_device_open": ([C.c_int32, C.POINTER(_P)], C.c_int),
    "ks_device_close": ([_P], C.c_int),
    "ks_device_set_timing": ([_P, C.c_int32], C.c_int),
    "ks_device_set_shading": ([_P, C.c_int32], C.c_int),
    "ks_device_set_tail": ([_P, C.c_int32], C.c_int),
    "ks_scene_create": ([_P, C.POINTER(SceneDesc), C.POINTER(BvhOptions), C.POINTER(_P)], C.c_int),
    "ks_scene_free": ([_P], C.c_int),
    "ks_scene_info_get": ([_P, C.POINTER(SceneInfo)], C.c_int),
    "ks_scene_light_table": ([_P, C.POINTER(C.c_int32), _P, _P, _P], C.c_int),
    "ks_scene_bvh_get": ([_P, _P, _P], C.c_int),
    "ks_film_create": ([_P, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(_P)], C.c_int),
    "ks_film_clear": ([_P], C.c_int),
    "ks_film_download": ([_P, _P, _P], C.c_int),
    "ks_film_composite": ([_P, _P], C.c_int),
    "ks_film_free": ([_P], C.c_int),
    "ks_film_relmse": ([_P, _P, C.POINTER(C.c_double)], C.c_int),
    "ks_film_relmse_noalloc": ([_P, _P, C.POINTER(C.c_double)], C.c_int),
    "ks_small_pt_render": (
        [_P, C.POINTER(CameraDesc), C.POINTER(SmallPTParams), _P, PROGRESS_FN, _P, C.POINTER(RenderStats)],
        C.c_int,
    ),
    "ks_sobol_stream": (
        [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _P, _P, _P, C.c_int32, _P, _P, _P],
        C.c_int,
    ),
    "ks_camera_rays": (
        [_P, C.POINTER(CameraDesc), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _P, _P, _P, _P],
        C.c_int,
    ),
    "ks_intersect": ([_P, _P, C.c_int64, _P], C.c_int),
    "ks_occluded": ([_P, _P, C.c_int64, _P], C.c_int),
    "ks_bsdf_eval": ([_P, C.POINTER(MaterialDesc), _P, _P, C.c_int64, _P, _P], C.c_int),
    "ks_bsdf_sample": ([_P, C.POINTER(MaterialDesc), _P, _P, C.c_int64, _P, _P, _P], C.c_int),
    "ks_trace_samples": (
        [_P, C.POINTER(CameraDesc), C.POINTER(SmallPTParams), C.c_int64, _P, _P, _P, _P],
        C.c_int,
    ),
    "ks_trace_capture": (
        [_P, C.POINTER(CameraDesc), C.POINTER(SmallPTParams), C.c_int64, _P, _P, _P, C.c_int64, _P,
         C.POINTER(C.c_int64), _P, C.POINTER(C.c_int64)],
        C.c_int,
    ),
    "ks_scene_traversal_counts": (
        [_P, _P, C.c_int64, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _P],
        C.c_int,
    ),
    "ks_small_pt_render_guided": (
        [_P, C.POINTER(CameraDesc), C.POINTER(SmallPTParams), _P, _P, PROGRESS_FN, _P, C.POINTER(RenderStats)],
        C.c_int,
    ),
    "ks_small_pt_render_batch": ([_P, C.c_int32, _P, C.POINTER(BatchOptions), C.POINTER(RenderStats)], C.c_int),
    "ks_guide_create": ([_P, C.POINTER(GuideConfig), _P, _P, C.POINTER(_P)], C.c_int),
    "ks_guide_free": ([_P], C.c_int),
    "ks_guide_set_mode": ([_P, C.c_float, C.c_int32], C.c_int),
    "ks_guide_get_stats": ([_P, C.POINTER(GuideStats)], C.c_int),
    "ks_guide_set_inference": ([_P, C.c_int32], C.c_int),
    "ks_guide_calibrate_fp8": ([_P, C.c_int64, _P], C.c_int),
    "ks_guide_fp8_scales": ([_P, _P], C.c_int),
    "ks_guide_param_count": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "ks_guide_grid_info": ([_P, _P, _P, _P], C.c_int),
    "ks_guide_get_params": ([_P, _P], C.c_int),
    "ks_guide_set_params": ([_P, _P], C.c_int),
    "ks_guide_forward": ([_P, C.c_int64, _P, _P], C.c_int),
    "ks_guide_gradient": ([_P, C.c_int64, _P, _P, _P, C.POINTER(C.c_double), _P], C.c_int),
    "ks_guide_train_step": ([_P, C.c_int64, _P, _P, _P, C.POINTER(C.c_double)], C.c_int),
    "ks_vmm_pdf": ([_P, C.c_int64, _P, _P, _P], C.c_int),
    "ks_vmm_sample": ([_P, C.c_int64, _P, _P, _P, _P], C.c_int),
    "ks_host_scene_create": ([C.c_char_p, C.c_uint64, C.POINTER(_P)], C.c_int),
    "ks_host_scene_get": ([_P, C.POINTER(SceneDesc), C.POINTER(CameraDesc)], C.c_int),
    "ks_host_scene_free": ([_P], C.c_int),
}
EXPORTED = tuple(_SIGS)

_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads libkeepsake_amd.so; raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is missing: build it with `python -m keepsake_amd.build`")
    lib = C.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _check(status: int) -> None:
    if status != KS_OK:
        msg = _lib.ks_last_error().decode(errors="replace") if _lib else "?"
        raise KsError(status, msg)


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    lib = load_library()
    n = C.c_int32(0)
    _check(lib.ks_device_count(C.byref(n)))
    return n.value


# ------------------------------------------------------------------------------------------------
# Scene description helpers
# ------------------------------------------------------------------------------------------------
MATERIAL_DTYPE = np.dtype(
    [
        ("basecolor", "<f4", 3),
        ("roughness", "<f4"),
        ("metallic", "<f4"),
        ("ior", "<f4"),
        ("specular_r0_mul", "<f4"),
        ("specular_trans", "<f4"),
        ("diffuse_trans", "<f4"),
        ("diffuse_trans_fwd", "<f4"),
        ("emission", "<f4", 3),
        ("has_emission", "<i4"),
    ]
)
LIGHT_DTYPE = np.dtype([("type", "<i4"), ("color", "<f4", 3), ("vec", "<f4", 3)])
assert MATERIAL_DTYPE.itemsize == C.sizeof(MaterialDesc)
assert LIGHT_DTYPE.itemsize == C.sizeof(LightDesc)


@dataclass
class HostScene:
    """Flat scene arrays (the ks_scene_desc contents) plus the camera of the scene."""

    vertices: np.ndarray  # (V, 3) float32
    indices: np.ndarray  # (T, 3) uint32
    tri_geom: np.ndarray  # (T,) uint32
    geom_material: np.ndarray  # (G,) uint32
    materials: np.ndarray  # MATERIAL_DTYPE
    lights: np.ndarray  # LIGHT_DTYPE
    camera: dict
    normals: np.ndarray | None = None
    geom_flags: np.ndarray | None = None  # (G,) uint8 of KS_GEOM_* bits, None: two-sided, smooth normals
    # environment map of a KS_LIGHT_SKY_MAP light: dict(rgb=(H, W, 3) float32, to_world=3x3, y_up=bool,
    # strength=float), or None
    sky_map: dict | None = None
    texcoords: np.ndarray | None = None  # (V, 2) float32 texture coordinates
    # textures: dicts with "data" ((H, W, C) uint8 / uint16 / float32, or a list of such arrays, one per
    # level), "srgb" (bool) and "mipmaps" (bool: build the chain from level 0)
    textures: list | None = None
    material_fields: dict | None = None  # material index -> MaterialFields
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_tris(self) -> int:
        return int(self.indices.shape[0])

    def desc(self) -> SceneDesc:
        arrs = [
            np.ascontiguousarray(self.vertices, np.float32),
            np.ascontiguousarray(self.indices, np.uint32),
            np.ascontiguousarray(self.tri_geom, np.uint32),
            np.ascontiguousarray(self.geom_material, np.uint32),
            np.ascontiguousarray(self.materials, MATERIAL_DTYPE),
            np.ascontiguousarray(self.lights, LIGHT_DTYPE),
        ]
        nrm = None if self.normals is None else np.ascontiguousarray(self.normals, np.float32)
        gf = None if self.geom_flags is None else np.ascontiguousarray(self.geom_flags, np.uint8)
        self._keep = arrs + [nrm, gf]
        v, idx, tg, gm, mats, lights = arrs
        d = SceneDesc()
        d.vertices = v.ctypes.data_as(C.POINTER(C.c_float))
        d.n_vertices = v.shape[0]
        d.normals = nrm.ctypes.data_as(C.POINTER(C.c_float)) if nrm is not None else None
        d.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
        d.n_tris = idx.shape[0]
        d.tri_geom = tg.ctypes.data_as(C.POINTER(C.c_uint32))
        d.geom_material = gm.ctypes.data_as(C.POINTER(C.c_uint32))
        d.n_geoms = gm.shape[0]
        d.materials = C.cast(mats.ctypes.data, C.POINTER(MaterialDesc))
        d.n_materials = mats.shape[0]
        d.lights = C.cast(lights.ctypes.data, C.POINTER(LightDesc)) if lights.shape[0] else None
        d.n_lights = lights.shape[0]
        d.geom_flags = gf.ctypes.data_as(C.POINTER(C.c_uint8)) if gf is not None else None
        if self.sky_map is not None:
            rgb = np.ascontiguousarray(self.sky_map["rgb"], np.float32)
            sm = SkyMapDesc()
            sm.rgb = rgb.ctypes.data_as(C.POINTER(C.c_float))
            sm.height, sm.width = int(rgb.shape[0]), int(rgb.shape[1])
            m = np.asarray(self.sky_map.get("to_world", np.eye(3)), np.float32).reshape(9)
            for k in range(9):
                sm.to_world[k] = float(m[k])
            sm.transform_y_up = 1 if self.sky_map.get("y_up", True) else 0
            sm.strength = float(self.sky_map.get("strength", 1.0))
            self._keep += [rgb, sm]
            d.sky_map = C.pointer(sm)
        if self.texcoords is not None:
            tc = np.ascontiguousarray(self.texcoords, np.float32)
            assert tc.shape == (v.shape[0], 2)
            self._keep.append(tc)
            d.texcoords = tc.ctypes.data_as(C.POINTER(C.c_float))
        if self.textures:
            arr = (TextureDesc * len(self.textures))()
            for i, t in enumerate(self.textures):
                levels = t["data"] if isinstance(t["data"], (list, tuple)) else [t["data"]]
                levels = [np.asarray(x) for x in levels]
                lv0 = levels[0] if levels[0].ndim == 3 else levels[0][..., None]
                blob = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(x).reshape(-1) for x in levels]))
                self._keep.append(blob)
                arr[i].data = blob.ctypes.data
                arr[i].height, arr[i].width, arr[i].channels = lv0.shape
                arr[i].data_type = _TEX_DTYPES[lv0.dtype]
                arr[i].color_space = KS_COLORSPACE_SRGB if t.get("srgb", False) else KS_COLORSPACE_LINEAR
                arr[i].levels = len(levels)
                arr[i].build_mipmaps = 1 if t.get("mipmaps", True) else 0
            self._keep.append(arr)
            d.textures = arr
            d.n_textures = len(self.textures)
        if self.material_fields:
            ptrs = (C.POINTER(MaterialFields) * mats.shape[0])()
            for m, mf in self.material_fields.items():
                self._keep.append(mf)
                ptrs[m] = C.pointer(mf)
            self._keep.append(ptrs)
            d.material_fields = ptrs
        return d


def camera_desc(cam: dict) -> CameraDesc:
    c = CameraDesc()
    if "to_world" in cam:
        c.use_matrix = 1
        m = np.asarray(cam["to_world"], np.float32).reshape(16)
        for i in range(16):
            c.to_world[i] = float(m[i])
    else:
        for i in range(3):
            c.pos[i] = float(cam["pos"][i])
            c.target[i] = float(cam["target"][i])
            c.up[i] = float(cam.get("up", (0.0, 1.0, 0.0))[i])
    c.vfov_deg = float(cam["vfov"])
    c.aspect = float(cam["aspect"])
    return c


def procedural_scene(name: str, seed: int = 0) -> HostScene:
    """Builds one of the benchmark scenes ("cornell", "spheres", "spheres_small", "room", "room_spheres")."""
    lib = load_library()
    h = C.c_void_p()
    _check(lib.ks_host_scene_create(name.encode(), seed, C.byref(h)))
    try:
        d = SceneDesc()
        cam = CameraDesc()
        _check(lib.ks_host_scene_get(h, C.byref(d), C.byref(cam)))

        def arr(ptr, n, dt):
            if n == 0 or not ptr:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_byte)), shape=(n * np.dtype(dt).itemsize,)).view(
                dt
            ).copy()

        verts = arr(d.vertices, d.n_vertices * 3, np.float32).reshape(-1, 3)
        normals = arr(d.normals, d.n_vertices * 3, np.float32).reshape(-1, 3) if d.normals else None
        idx = arr(d.indices, d.n_tris * 3, np.uint32).reshape(-1, 3)
        tg = arr(d.tri_geom, d.n_tris, np.uint32)
        gm = arr(d.geom_material, d.n_geoms, np.uint32)
        mats = arr(d.materials, d.n_materials, MATERIAL_DTYPE)
        lights = arr(d.lights, d.n_lights, LIGHT_DTYPE)
        camera = {
            "pos": tuple(cam.pos),
            "target": tuple(cam.target),
            "up": tuple(cam.up),
            "vfov": cam.vfov_deg,
            "aspect": cam.aspect,
        }
        return HostScene(verts, idx, tg, gm, mats, lights, camera, normals)
    finally:
        lib.ks_host_scene_free(h)


def small_pt_params(**kw) -> SmallPTParams:
    """SmallPTInput defaults (small_pt.h:17-41) with keyword overrides."""
    p = SmallPTParams()
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
    p.scale_ray_diff = int(kw.get("scale_ray_diff", True))
    p.clamp_direct = kw.get("clamp_direct", 0.0)
    p.clamp_indirect = kw.get("clamp_indirect", 10.0)
    p.spp_prog_interval = kw.get("spp_prog_interval", 32)
    p.mis_mode = kw.get("mis_mode", KS_MIS_REFERENCE)
    p.guide_mode = kw.get("guide_mode", KS_GUIDE_OFF)
    p.queue_size = kw.get("queue_size", 0)
    p.pixel_filter = kw.get("pixel_filter", KS_FILTER_BOX)
    p.filter_width = kw.get("filter_width", 0.0)
    p.filter_sigma = kw.get("filter_sigma", 0.0)
    return p


# ------------------------------------------------------------------------------------------------
# Handles
# ------------------------------------------------------------------------------------------------
class Device:
    def __init__(self, ordinal: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.ks_device_open(ordinal, C.byref(h)))
        self.h = h

    def set_timing(self, enable: bool) -> None:
        _check(self.lib.ks_device_set_timing(self.h, int(enable)))

    def set_shading(self, sorted: bool) -> None:
        """Material-sorted shading, or one shading kernel over every path (the default)."""
        _check(self.lib.ks_device_set_shading(self.h, int(sorted)))

    def set_tail(self, max_paths: int) -> None:
        """The render's last max_paths paths (or fewer) go to the tail kernel (0: off)."""
        _check(self.lib.ks_device_set_tail(self.h, int(max_paths)))

    def close(self) -> None:
        if self.h:
            self.lib.ks_device_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- component entry points -----------------------------------------------------------------
    def sobol_stream(self, width, height, spp, seed, px, py, sample, kinds):
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        kinds = np.ascontiguousarray(kinds, np.int32)
        n = px.shape[0]
        nv = int(kinds.sum())
        ou = np.zeros((n, nv), np.uint32)
        of = np.zeros((n, nv), np.float32)
        _check(
            self.lib.ks_sobol_stream(
                self.h, width, height, spp, seed, n, _ptr(px), _ptr(py), _ptr(sample), kinds.shape[0], _ptr(kinds),
                _ptr(ou), _ptr(of),
            )
        )
        return ou, of

    def camera_rays(self, cam: dict, width, height, spp, seed, px, py, sample):
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        n = px.shape[0]
        out = np.zeros((n, 6), np.float32)
        c = camera_desc(cam)
        _check(
            self.lib.ks_camera_rays(
                self.h, C.byref(c), width, height, spp, seed, n, _ptr(px), _ptr(py), _ptr(sample), _ptr(out)
            )
        )
        return out

    def bsdf_eval(self, material, wo, wi):
        m = _material_struct(material)
        wo = np.ascontiguousarray(wo, np.float32)
        wi = np.ascontiguousarray(wi, np.float32)
        n = wo.shape[0]
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        _check(self.lib.ks_bsdf_eval(self.h, C.byref(m), _ptr(wo), _ptr(wi), n, _ptr(f), _ptr(pdf)))
        return f, pdf

    def bsdf_sample(self, material, wo, u):
        m = _material_struct(material)
        wo = np.ascontiguousarray(wo, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        n = wo.shape[0]
        wi = np.zeros((n, 3), np.float32)
        beta = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        _check(self.lib.ks_bsdf_sample(self.h, C.byref(m), _ptr(wo), _ptr(u), n, _ptr(wi), _ptr(beta), _ptr(pdf)))
        return wi, beta, pdf


def _material_struct(m) -> MaterialDesc:
    if isinstance(m, MaterialDesc):
        return m
    rec = np.asarray(m, MATERIAL_DTYPE).reshape(())
    s = MaterialDesc()
    C.memmove(C.byref(s), rec.tobytes(), C.sizeof(s))
    return s


class Scene:
    """Uploaded scene: geometry, SBVH and light tables on the device (Scene + PowerLightSampler)."""

    def __init__(self, device: Device, hs: HostScene, bvh: dict | None = None):
        self.device = device
        self.lib = device.lib
        self.host = hs
        d = hs.desc()
        o = BvhOptions()
        for k, v in (bvh or {}).items():
            setattr(o, k, v)
        h = C.c_void_p()
        _check(self.lib.ks_scene_create(device.h, C.byref(d), C.byref(o), C.byref(h)))
        self.h = h

    def info(self) -> dict:
        i = SceneInfo()
        _check(self.lib.ks_scene_info_get(self.h, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in i._fields_}
        d["bound_lo"] = tuple(i.bound_lo)
        d["bound_hi"] = tuple(i.bound_hi)
        return d

    def bvh(self):
        """The device BVH: (nodes, refs) as uint32 arrays of 32 (BVH4) or 16 (BVH2) words per node and
        12 words per triangle record."""
        i = self.info()
        nodes = np.zeros((i["n_nodes"], 32 if i["branching"] == 4 else 16), np.uint32)
        refs = np.zeros((i["n_refs"], 12), np.uint32)
        _check(self.lib.ks_scene_bvh_get(self.h, _ptr(nodes), _ptr(refs)))
        return nodes, refs

    def light_table(self):
        n = C.c_int32(0)
        _check(self.lib.ks_scene_light_table(self.h, C.byref(n), None, None, None))
        prim = np.zeros(n.value, np.int32)
        power = np.zeros(n.value, np.float32)
        cdf = np.zeros(n.value + 1, np.float32)
        _check(self.lib.ks_scene_light_table(self.h, C.byref(n), _ptr(prim), _ptr(power), _ptr(cdf)))
        return prim, power, cdf

    def intersect(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((rays.shape[0], 4), np.float32)
        _check(self.lib.ks_intersect(self.h, _ptr(rays), rays.shape[0], _ptr(out)))
        return out

    def occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(rays.shape[0], np.uint8)
        _check(self.lib.ks_occluded(self.h, _ptr(rays), rays.shape[0], _ptr(out)))
        return out

    def trace_samples(self, cam: dict, params: SmallPTParams, px, py, sample) -> np.ndarray:
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        out = np.zeros((px.shape[0], 4), np.float32)
        c = camera_desc(cam)
        _check(
            self.lib.ks_trace_samples(
                self.h, C.byref(c), C.byref(params), px.shape[0], _ptr(px), _ptr(py), _ptr(sample), _ptr(out)
            )
        )
        return out

    def capture_rays(self, cam: dict, params: SmallPTParams, px, py, sample, max_rays: int):
        """Every closest-hit and shadow ray the listed paths issue: two (m, 8) arrays (o, tmin, d, tmax)."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        rc = np.zeros((max_rays, 8), np.float32)
        rs = np.zeros((max_rays, 8), np.float32)
        nc, ns = C.c_int64(), C.c_int64()
        c = camera_desc(cam)
        _check(self.lib.ks_trace_capture(self.h, C.byref(c), C.byref(params), px.shape[0], _ptr(px), _ptr(py),
                                         _ptr(sample), max_rays, _ptr(rc), C.byref(nc), _ptr(rs), C.byref(ns)))
        return rc[: nc.value], rs[: ns.value]

    def traversal_counts(self, rays, any_hit: bool = False, with_t: bool = False):
        """Node fetches and triangle tests of the reference traversal order (host SBVH, BVH2) over rays."""
        rays = np.ascontiguousarray(rays, np.float32)
        nodes, tris = C.c_uint64(), C.c_uint64()
        t = np.zeros(rays.shape[0], np.float32) if with_t else None
        _check(self.lib.ks_scene_traversal_counts(self.h, _ptr(rays), rays.shape[0], int(any_hit), C.byref(nodes),
                                                  C.byref(tris), _ptr(t) if with_t else None))
        return (nodes.value, tris.value, t) if with_t else (nodes.value, tris.value)

    def close(self) -> None:
        if self.h:
            self.lib.ks_scene_free(self.h)
            self.h = None


class Film:
    """RenderTarget2: FP64 RGB sums + (hit, miss) weights on the device."""

    def __init__(self, device: Device, width: int, height: int, backdrop=(0.0, 0.0, 0.0)):
        self.device = device
        self.lib = device.lib
        self.width, self.height = width, height
        bd = (C.c_float * 3)(*backdrop)
        h = C.c_void_p()
        _check(self.lib.ks_film_create(device.h, width, height, bd, C.byref(h)))
        self.h = h

    def clear(self) -> None:
        _check(self.lib.ks_film_clear(self.h))

    def download(self):
        rgb = np.zeros((self.height, self.width, 3), np.float64)
        w = np.zeros((self.height, self.width, 2), np.float64)
        _check(self.lib.ks_film_download(self.h, _ptr(rgb), _ptr(w)))
        return rgb, w

    def composite(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float32)
        _check(self.lib.ks_film_composite(self.h, _ptr(out)))
        return out

    def relmse_noalloc(self, reference: "Film") -> float:
        """relmse() through the allocation-free entry point (static device scratch)."""
        v = C.c_double()
        _check(self.lib.ks_film_relmse_noalloc(self.h, reference.h, C.byref(v)))
        return v.value

    def relmse(self, reference: "Film") -> float:
        """mean over pixels and channels of (x - r)^2 / (r^2 + 0.01) of the composites, FP64 on the device"""
        v = C.c_double()
        _check(self.lib.ks_film_relmse(self.h, reference.h, C.byref(v)))
        return v.value

    def close(self) -> None:
        if self.h:
            self.lib.ks_film_free(self.h)
            self.h = None


def small_pt_render(scene: Scene, cam: dict, params: SmallPTParams, film: Film, progress=None) -> dict:
    """SmallPT::run: renders into `film`; returns the render statistics."""
    st = RenderStats()
    c = camera_desc(cam)
    if progress is not None:
        cb = PROGRESS_FN(lambda f, s, u: progress(film, s))
    else:
        cb = PROGRESS_FN()
    _check(scene.lib.ks_small_pt_render(scene.h, C.byref(c), C.byref(params), film.h, cb, None, C.byref(st)))
    return st.as_dict()


def small_pt_render_guided(scene: Scene, cam: dict, params: SmallPTParams, guide: "Guide", film: Film,
                           progress=None) -> dict:
    """small_pt with the learned guide (one-sample MIS, online training between 1-spp passes)."""
    st = RenderStats()
    c = camera_desc(cam)
    cb = PROGRESS_FN(lambda f, s, u: progress(film, s)) if progress is not None else PROGRESS_FN()
    _check(scene.lib.ks_small_pt_render_guided(scene.h, C.byref(c), C.byref(params), guide.h, film.h, cb, None,
                                               C.byref(st)))
    return st.as_dict()


def small_pt_render_batch(scene: Scene, views, queue_size: int = 0, trace_blocks_per_cu: int = 0,
                          graph: bool = True, max_chunk_items: int = 0) -> dict:
    """Many views of one scene in one wavefront: views is a list of (camera dict, SmallPTParams, Film)."""
    arr = (BatchView * len(views))()
    for i, (cam, params, film) in enumerate(views):
        arr[i].camera = camera_desc(cam)
        arr[i].params = params
        arr[i].film = film.h
    o = BatchOptions(queue_size, trace_blocks_per_cu, 0 if graph else 1, max_chunk_items)
    st = RenderStats()
    _check(scene.lib.ks_small_pt_render_batch(scene.h, len(views), C.cast(arr, C.c_void_p), C.byref(o),
                                              C.byref(st)))
    return st.as_dict()


def guide_config(**kw) -> GuideConfig:
    c = GuideConfig()
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Guide:
    """The learned path guide (hash grid + MLP -> 8-lobe vMF mixture) on one device."""

    def __init__(self, device: Device, bound_lo, bound_hi, config: GuideConfig | None = None):
        self.device = device
        self.lib = device.lib
        lo = np.ascontiguousarray(bound_lo, np.float32)
        hi = np.ascontiguousarray(bound_hi, np.float32)
        h = C.c_void_p()
        cfg = config if config is not None else GuideConfig()
        _check(self.lib.ks_guide_create(device.h, C.byref(cfg), _ptr(lo), _ptr(hi), C.byref(h)))
        self.h = h
        ng, nm = C.c_int64(), C.c_int64()
        _check(self.lib.ks_guide_param_count(self.h, C.byref(ng), C.byref(nm)))
        self.n_grid, self.n_mlp = ng.value, nm.value

    def close(self) -> None:
        if self.h:
            self.lib.ks_guide_free(self.h)
            self.h = None

    def set_mode(self, guide_fraction: float, max_steps_per_pass: int) -> None:
        _check(self.lib.ks_guide_set_mode(self.h, guide_fraction, max_steps_per_pass))

    def set_inference(self, fp8: bool) -> None:
        """FP8 (e4m3) or BF16 MLP inference for later forwards and guided renders."""
        _check(self.lib.ks_guide_set_inference(self.h, 1 if fp8 else 0))

    def calibrate_fp8(self, queries) -> None:
        """Record each layer's largest |input| over the queries and rebuild the FP8 scales from them."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        _check(self.lib.ks_guide_calibrate_fp8(self.h, q.shape[0], _ptr(q)))

    def fp8_scales(self) -> dict:
        v = np.zeros(12, np.float32)
        _check(self.lib.ks_guide_fp8_scales(self.h, _ptr(v)))
        names = ["inv_sx", "inv_sh1", "inv_sh2", "c1", "c2", "c3", "sw1", "sw2", "sw3", "sx", "sh1", "sh2"]
        return {k: float(x) for k, x in zip(names, v)}

    def stats(self) -> dict:
        st = GuideStats()
        _check(self.lib.ks_guide_get_stats(self.h, C.byref(st)))
        d = {k: getattr(st, k) for k, _ in st._fields_}
        w = d["last_weight"]
        d["cross_entropy"] = d["last_loss"] / w if w > 0 else float("nan")
        d["cross_entropy_cosine"] = d["last_loss_cosine"] / w if w > 0 else float("nan")
        return d

    def grid_info(self):
        res = np.zeros(16, np.int32)
        lo = np.zeros(3, np.float32)
        inv = np.zeros(3, np.float32)
        _check(self.lib.ks_guide_grid_info(self.h, _ptr(res), _ptr(lo), _ptr(inv)))
        return res, lo, inv

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_grid + self.n_mlp, np.float32)
        _check(self.lib.ks_guide_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_grid + self.n_mlp:
            raise ValueError("parameter vector has the wrong size")
        _check(self.lib.ks_guide_set_params(self.h, _ptr(p)))

    def forward(self, queries) -> np.ndarray:
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        out = np.zeros((q.shape[0], 40), np.float32)
        _check(self.lib.ks_guide_forward(self.h, q.shape[0], _ptr(q), _ptr(out)))
        return out

    def _samples(self, queries, wi, weight):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        w = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
        c = np.ascontiguousarray(weight, np.float32).reshape(-1)
        if not (len(q) == len(w) == len(c)):
            raise ValueError("queries, wi and weight differ in length")
        return q, w, c

    def gradient(self, queries, wi, weight):
        q, w, c = self._samples(queries, wi, weight)
        g = np.zeros(self.n_grid + self.n_mlp, np.float32)
        loss = C.c_double()
        _check(self.lib.ks_guide_gradient(self.h, len(q), _ptr(q), _ptr(w), _ptr(c), C.byref(loss), _ptr(g)))
        return loss.value, g

    def train_step(self, queries, wi, weight) -> float:
        q, w, c = self._samples(queries, wi, weight)
        loss = C.c_double()
        _check(self.lib.ks_guide_train_step(self.h, len(q), _ptr(q), _ptr(w), _ptr(c), C.byref(loss)))
        return loss.value


def vmm_pdf(device: Device, mixtures, dirs) -> np.ndarray:
    m = np.ascontiguousarray(mixtures, np.float32).reshape(-1, 40)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    p = np.zeros(len(m), np.float32)
    _check(device.lib.ks_vmm_pdf(device.h, len(m), _ptr(m), _ptr(d), _ptr(p)))
    return p


def vmm_sample(device: Device, mixtures, u):
    m = np.ascontiguousarray(mixtures, np.float32).reshape(-1, 40)
    uu = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    d = np.zeros((len(m), 3), np.float32)
    p = np.zeros(len(m), np.float32)
    _check(device.lib.ks_vmm_sample(device.h, len(m), _ptr(m), _ptr(uu), _ptr(d), _ptr(p)))
    return d, p
