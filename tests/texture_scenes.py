"""Textured test scenes built in Python (HostScene with texcoords, textures and material fields).

textured_box(): a unit box lit by a ceiling quad whose emission is an HDR f32 texture (EWA sampler),
with
  - a floor with a tiled u8 sRGB checker basecolor (linear sampler, mip chain, repeat wrap, uv x4),
  - a back wall with a u16 tangent-space normal map (bilinear reads) over a u8 RGBA basecolor whose
    alpha drives nothing (swizzle picks RGB),
  - a right wall whose roughness (f32, Catmull-Rom cubic) and metallic (u8 linear, nearest, clamp) are
    textures,
  - a left wall with no texture coordinates (KS_GEOM_NO_TEXCOORD) and a basecolor texture read at the
    barycentric (u, v), cubic B-spline,
  - a free-standing quad with a u8 cut-out opacity map (nearest) and one with a constant opacity 0.5
    (or 1),
  - a short box of constant material (uv-less faces of a textured scene).
"""

import numpy as np

from keepsake_amd import api


def checker(n, cells, c0, c1):
    y, x = np.mgrid[0:n, 0:n]
    m = ((x * cells // n) + (y * cells // n)) % 2 == 0
    return np.where(m[..., None], np.asarray(c0, np.float64), np.asarray(c1, np.float64))


class Builder:
    def __init__(self):
        self.v, self.tc, self.idx, self.tg, self.gm, self.flags = [], [], [], [], [], []
        self.n_geoms = 0

    def quad(self, p0, p1, p2, p3, uv=((0, 0), (1, 0), (1, 1), (0, 1))):
        b = len(self.v)
        self.v += [p0, p1, p2, p3]
        self.tc += list(uv)
        self.idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        self.tg += [self.n_geoms, self.n_geoms]

    def box(self, lo, hi):
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        self.quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1))
        self.quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1))
        self.quad((x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0))
        self.quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0))
        self.quad((x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1))

    def end(self, material, flags=0):
        self.gm.append(material)
        self.flags.append(flags)
        self.n_geoms += 1


def textured_box(emission_scale=4.0, stochastic_opacity=True):
    """stochastic_opacity=False gives the second quad an opacity of 1 instead of 0.5: then every decision
    is a deterministic function of the hit, and per-sample radiance can be compared across
    implementations whose ray floats differ in the last bit (the stochastic test hashes those bits)."""
    b = Builder()
    b.quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), uv=((0, 0), (0, 4), (4, 4), (4, 0)))  # floor, tiled
    b.end(0)
    b.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))  # back wall: normal map
    b.end(1)
    b.quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))  # right wall: roughness / metallic textures
    b.end(2)
    b.quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1))  # left wall: no texcoords
    b.end(3, api.KS_GEOM_NO_TEXCOORD)
    b.quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0))  # ceiling
    b.end(4)
    y = 0.998
    b.quad((0.3, y, 0.3), (0.7, y, 0.3), (0.7, y, 0.7), (0.3, y, 0.7))  # textured emitter
    b.end(5)
    b.quad((0.15, 0.05, 0.55), (0.55, 0.05, 0.55), (0.55, 0.55, 0.55), (0.15, 0.55, 0.55))  # cut-out
    b.end(6)
    b.quad((0.5, 0.1, 0.75), (0.9, 0.1, 0.75), (0.9, 0.45, 0.75), (0.5, 0.45, 0.75))  # half-transparent
    b.end(7)
    b.box((0.6, 0.0, 0.2), (0.85, 0.3, 0.45))  # constant-material box
    b.end(8)

    mats = np.zeros(9, api.MATERIAL_DTYPE)
    mats["basecolor"] = 0.7
    mats["roughness"] = 0.6
    mats["ior"] = 1.5
    mats["specular_r0_mul"] = 0.5
    mats[3]["basecolor"] = (0.7, 0.1, 0.1)
    mats[5]["basecolor"] = 0.0
    mats[5]["has_emission"] = 1

    rng = np.random.default_rng(5)
    tex = []
    # 0: floor checker, u8 sRGB RGB, 100 x 100 (non power of two chain)
    c = checker(100, 10, (0.9, 0.85, 0.8), (0.1, 0.15, 0.3))
    tex.append(dict(data=np.round(c * 255).astype(np.uint8), srgb=True, mipmaps=True))
    # 1: normal map, u16 linear RGB 64 x 48: bumps in the tangent plane, encoded in [0, 1]
    yy, xx = np.mgrid[0:48, 0:64]
    nx = 0.35 * np.sin(xx / 64 * 6 * np.pi)
    ny = 0.25 * np.cos(yy / 48 * 4 * np.pi)
    n = np.stack([nx, ny, np.ones_like(nx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    tex.append(dict(data=np.round((n * 0.5 + 0.5) * 65535).astype(np.uint16), srgb=False, mipmaps=False))
    # 2: back-wall basecolor, u8 RGBA sRGB 33 x 17
    bc = rng.integers(40, 230, (17, 33, 4)).astype(np.uint8)
    tex.append(dict(data=bc, srgb=True, mipmaps=True))
    # 3: roughness, f32 one channel 20 x 30
    r = (0.05 + 0.9 * (np.mgrid[0:30, 0:20][1] / 19.0) ** 2).astype(np.float32)[..., None]
    tex.append(dict(data=r, srgb=False, mipmaps=True))
    # 4: metallic, u8 linear one channel 8 x 8 stripes
    met = np.where((np.arange(8) // 2) % 2 == 0, 230, 10).astype(np.uint8)[None, :, None].repeat(8, axis=0)
    tex.append(dict(data=met, srgb=False, mipmaps=False))
    # 5: left-wall basecolor at barycentrics, u16 sRGB RGB 16 x 16
    lw = rng.integers(0, 65535, (16, 16, 3)).astype(np.uint16)
    tex.append(dict(data=lw, srgb=True, mipmaps=True))
    # 6: emission, f32 RGB 24 x 24 HDR pattern
    ey, ex = np.mgrid[0:24, 0:24]
    em = np.stack([8 + 6 * np.sin(ex / 3.0), 6 + 5 * np.cos(ey / 4.0), 4 + 3 * ((ex + ey) % 5 == 0)], -1)
    tex.append(dict(data=em.astype(np.float32), srgb=False, mipmaps=True))
    # 7: cut-out mask, u8 one channel 12 x 12: a ring of holes
    my, mx = np.mgrid[0:12, 0:12]
    mask = np.where(((mx - 5.5) ** 2 + (my - 5.5) ** 2) < 12, 0, 255).astype(np.uint8)[..., None]
    tex.append(dict(data=mask, srgb=False, mipmaps=False))

    F = api.field_desc
    fields = {
        0: api.material_fields(mats[0], basecolor=F(texture=0, sampler=api.KS_SAMPLER_LINEAR)),
        1: api.material_fields(mats[1], basecolor=F(texture=2, sampler=api.KS_SAMPLER_LINEAR, swizzle=(0, 1, 2)),
                               normal_map=F(texture=1, sampler=api.KS_SAMPLER_LINEAR), normal_strength=0.8),
        2: api.material_fields(mats[2], roughness=F(texture=3, sampler=api.KS_SAMPLER_CUBIC,
                                                    cubic_kernel=api.KS_CUBIC_CATMULL_ROM),
                               metallic=F(texture=4, sampler=api.KS_SAMPLER_NEAREST, wrap_u=api.KS_WRAP_CLAMP,
                                          wrap_v=api.KS_WRAP_CLAMP, swizzle=(0,)),
                               basecolor=F(value=(0.8, 0.6, 0.3))),
        3: api.material_fields(mats[3], basecolor=F(texture=5, sampler=api.KS_SAMPLER_CUBIC,
                                                    cubic_kernel=api.KS_CUBIC_BSPLINE, scale=(1.0, 0.8, 0.6))),
        5: api.material_fields(mats[5], emission=F(texture=6, sampler=api.KS_SAMPLER_EWA, anisotropy=4.0,
                                                   scale=(emission_scale / 8,) * 3)),
        6: api.material_fields(mats[6], basecolor=F(value=(0.2, 0.6, 0.3)),
                               opacity=F(texture=7, sampler=api.KS_SAMPLER_NEAREST, swizzle=(0,))),
        7: api.material_fields(mats[7], basecolor=F(value=(0.3, 0.3, 0.8)),
                               opacity=F(value=0.5 if stochastic_opacity else 1.0)),
    }
    hs = api.HostScene(
        vertices=np.asarray(b.v, np.float32), indices=np.asarray(b.idx, np.uint32),
        tri_geom=np.asarray(b.tg, np.uint32), geom_material=np.asarray(b.gm, np.uint32), materials=mats,
        lights=np.zeros(0, api.LIGHT_DTYPE),
        camera=dict(pos=(0.5, 0.5, 2.2), target=(0.5, 0.45, 0.4), up=(0.0, 1.0, 0.0), vfov=42.0, aspect=1.0),
        geom_flags=np.asarray(b.flags, np.uint8), texcoords=np.asarray(b.tc, np.float32), textures=tex,
        material_fields=fields)
    return hs
