"""Texture assets of the TOML path (keepsake_amd/host/texture_asset.cpp, image_io.cpp), checked through
`ks_render --dump-scene` (no GPU):

- PNG input as stb_image delivers it to create_texture_from_image (texture.cpp:922-948): every colour
  type and bit depth, all five row filters, Adam7, palettes and tRNS, 16-bit samples cut to 8, and the
  requested channel count converted by stb's rules; EXR / HDR inputs as f32 with their channel rules;
- the serialized texture format (texture.cpp:950-1066): written by convert_texture with the full mip
  chain (equal to the oracle's chain) in LZ4 blocks that an independent decoder reads, and read back as a
  linear texture;
- shader-field and normal-map tables (shader_field.h:88-158, texture.cpp:1086-1121, normal_map.cpp:42-56)
  into the ks_material_fields records, texture coordinates of OBJ meshes, and the errors.
"""

import ctypes as C
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import image_files as imf
from keepsake_amd import api

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "keepsake_amd" / "_lib" / "ks_render"

QUAD_OBJ = """v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
f 1/1 2/2 3/3 4/4
"""


def config(tmp_path, textures, bsdf_extra="", material_extra="", extra_assets=""):
    (tmp_path / "quad.obj").write_text(QUAD_OBJ)
    tex = "".join(f"[texture.{n}]\n{body}\n" for n, body in textures.items())
    return f"""
output_dir = "{tmp_path / 'out'}"
{tex}
{extra_assets}
[mesh_asset.quad]
path = "{tmp_path / 'quad.obj'}"
format = "obj"
load_materials = false

[bsdf.b]
type = "principled_bsdf"
roughness = {{ type = "constant", value = 0.5 }}
metallic = {{ type = "constant", value = 0.0 }}
ior = {{ type = "constant", value = 1.5 }}
specular_r0_mul = {{ type = "constant", value = 0.5 }}
specular_trans = {{ type = "constant", value = 0.0 }}
diffuse_trans = {{ type = "constant", value = 0.0 }}
diffuse_trans_fwd = {{ type = "constant", value = 0.0 }}
{bsdf_extra}

[material.m]
bsdf = "bsdf.b"
{material_extra}

[[task]]
type = "small_pt"
object = "mesh_asset.quad"
material = [["material.m"]]
bounces = 2
render_width = 16
render_height = 16
spp = 1
[task.camera]
type = "perspective"
pos = [0.5, 0.5, 2.0]
target = [0.5, 0.5, 0.0]
up = [0.0, 1.0, 0.0]
vfov = 40.0
aspect = 1.0
"""


def dump(tmp_path, text):
    cfg = tmp_path / "c.toml"
    cfg.write_text(text)
    out = tmp_path / "dump"
    r = subprocess.run([str(TOOL), "-c", str(cfg), "--dump-scene", str(out)], capture_output=True, text=True)
    return r, out / "task_0"


def read_dump(d):
    res = {"texcoords": np.fromfile(d / "texcoords.f32", np.float32).reshape(-1, 2),
           "field_of": np.fromfile(d / "field_of.i32", np.int32)}
    raw = (d / "fields.bin").read_bytes()
    n = len(raw) // C.sizeof(api.MaterialFields)
    res["fields"] = [api.MaterialFields.from_buffer_copy(raw, i * C.sizeof(api.MaterialFields)) for i in range(n)]
    res["textures"] = []
    for i, line in enumerate((d / "textures.txt").read_text().split("\n")):
        if line.strip():
            w, h, c, dt, cs, lv, mip = map(int, line.split())
            res["textures"].append(dict(w=w, h=h, c=c, dtype=dt, cs=cs, levels=lv, mips=mip,
                                        bytes=(d / f"texture_{i}.bin").read_bytes()))
    return res


BASECOLOR_TEX = 'basecolor = { type = "texture", map = "texture.t", sampler = { type = "linear" } }'


# ---- PNG decoding -----------------------------------------------------------------------------------

def stb_expand(samples, ctype, depth, palette, trns):
    """What stb_image delivers before the channel request: (H, W, comp) and the sample maximum."""
    s = samples.astype(np.int64)
    full = 65535 if depth == 16 else 255
    if ctype == 3:
        pal = np.asarray(palette, np.int64)
        rgb = pal[s[..., 0]]
        if trns is not None:
            a = np.array(list(trns) + [255] * (len(pal) - len(trns)), np.int64)[s[..., 0]]
            return np.concatenate([rgb, a[..., None]], -1), 255
        return rgb, 255
    if depth < 8:
        s = s * {1: 255, 2: 85, 4: 17}[depth]
    if trns is not None and ctype in (0, 2):
        c = s.shape[-1]
        key = np.array([trns[2 * k] << 8 | trns[2 * k + 1] for k in range(c)]) & ((1 << depth) - 1)
        raw = samples.astype(np.int64)
        a = np.where(np.all(raw == key, axis=-1), 0, full)
        s = np.concatenate([s, a[..., None]], -1)
    return s, full


def stb_channels(img, full, req):
    comp = img.shape[-1]
    if comp == req:
        return img
    if comp >= 3:
        g = (img[..., 0] * 77 + img[..., 1] * 150 + img[..., 2] * 29) >> 8 if full == 255 else \
            (img[..., 0] * 19595 + img[..., 1] * 38470 + img[..., 2] * 7471) >> 16
    else:
        g = img[..., 0]
    a = img[..., 1] if comp == 2 else (img[..., 3] if comp == 4 else np.full(g.shape, full))
    rgb = img[..., :3] if comp >= 3 else np.stack([g, g, g], -1)
    return {1: g[..., None], 2: np.stack([g, a], -1), 3: rgb, 4: np.concatenate([rgb, a[..., None]], -1)}[req]


PNG_CASES = [
    (0, 1, False, False), (0, 2, False, True), (0, 4, True, False), (0, 8, False, False), (0, 16, True, False),
    (2, 8, False, False), (2, 16, False, True), (2, 8, True, False), (3, 1, False, False), (3, 2, True, True),
    (3, 4, False, False), (3, 8, True, True), (4, 8, False, False), (4, 16, True, False), (6, 8, False, True),
    (6, 16, False, False),
]


@pytest.mark.parametrize("ctype,depth,trns,interlace", PNG_CASES)
@pytest.mark.parametrize("req", [1, 3, 4])
def test_png_textures(tmp_path, ctype, depth, trns, interlace, req):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * trns)
    h, w = 11, 13
    c = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    if ctype == 3:
        palette = rng.integers(0, 256, (1 << depth, 3))
        samples = rng.integers(0, 1 << depth, (h, w, 1))
        t = bytes(rng.integers(0, 256, max(1, (1 << depth) // 2)).astype(np.uint8)) if trns else None
    else:
        palette = None
        samples = rng.integers(0, top + 1, (h, w, c))
        t = None
        if trns and ctype in (0, 2):
            key = samples[3, 4]  # a colour that occurs
            t = b"".join(struct.pack(">H", int(k)) for k in key)
    path = tmp_path / "t.png"
    imf.write_png_any(path, samples, ctype, depth, palette=palette, trns=t, interlace=interlace)
    exp, full = stb_expand(samples, ctype, depth, palette, t)
    if full == 65535:
        exp = exp >> 8  # stbi_load: 16 -> 8 bits by the high byte
        full = 255
    exp = stb_channels(exp, full, req).astype(np.uint8)
    r, d = dump(tmp_path, config(tmp_path, {"t": f'path = "{path}"\nchannels = {req}\nbuild_mipmaps = false'},
                                 bsdf_extra=BASECOLOR_TEX.replace('linear" }', 'linear" }, swizzle = [0, 0, 0]')))
    assert r.returncode == 0, r.stderr
    t0 = read_dump(d)["textures"][0]
    assert (t0["w"], t0["h"], t0["c"], t0["dtype"]) == (w, h, req, api.KS_TEX_U8)
    got = np.frombuffer(t0["bytes"], np.uint8).reshape(h, w, req)
    np.testing.assert_array_equal(got, exp)


def test_exr_and_hdr_textures(tmp_path):
    rng = np.random.default_rng(3)
    rgba = rng.uniform(0, 4, (5, 6, 4)).astype(np.float32)
    imf.write_exr(tmp_path / "a.exr", rgba, channels="RGBA")
    imf.write_exr(tmp_path / "y.exr", rgba[..., :1], channels="Y")
    rgbe = imf.to_rgbe(rgba[..., :3])
    imf.write_hdr(tmp_path / "h.hdr", rgbe, rle=False)  # run-length scanlines need widths of 8 or more
    hdr = imf.rgbe_to_float(rgbe)
    cases = {"a": ("a.exr", 4, rgba), "b": ("a.exr", 2, rgba[..., :2]), "y": ("y.exr", 4, np.repeat(rgba[..., :1], 4, -1)),
             "h1": ("h.hdr", 1, hdr.mean(-1, keepdims=True)), "h4": ("h.hdr", 4, np.concatenate([hdr, np.ones_like(hdr[..., :1])], -1))}
    for name, (f, ch, want) in cases.items():
        text = config(tmp_path, {"t": f'path = "{tmp_path / f}"\nchannels = {ch}\nbuild_mipmaps = true'},
                      bsdf_extra=BASECOLOR_TEX.replace('linear" }', 'linear" }, swizzle = [0, 0, 0]'))
        r, d = dump(tmp_path, text)
        assert r.returncode == 0, (name, r.stderr)
        t0 = read_dump(d)["textures"][0]
        assert (t0["c"], t0["dtype"], t0["levels"], t0["mips"]) == (ch, api.KS_TEX_F32, 1, 1)
        got = np.frombuffer(t0["bytes"], np.float32).reshape(5, 6, ch)
        np.testing.assert_allclose(got, want, rtol=1e-6, err_msg=name)


# ---- serialized textures ------------------------------------------------------------------------------

def parse_serialized(path):
    b = path.read_bytes()
    assert b[:25] == b"i_am_a_serialized_texture"
    w, h, c, dt, levels = struct.unpack_from("<5i", b, 25)
    (total,) = struct.unpack_from("<Q", b, 45)
    (nb,) = struct.unpack_from("<i", b, 53)
    sizes = struct.unpack_from(f"<{nb}i", b, 57)
    p = 57 + 4 * nb
    data = b""
    for s in sizes:
        data += imf.lz4_decode(b[p:p + s], min(total - len(data), 0x7E000000))
        p += s
    assert len(data) == total and p == len(b)
    return w, h, c, dt, levels, data


def test_convert_and_read_serialized_texture(tmp_path, ora):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (37, 21, 3)).astype(np.uint8)
    img[:, :10] = img[0, 0]  # repeated runs for the LZ4 matcher
    imf.write_png_any(tmp_path / "s.png", img, 2, 8)
    text = f"""
output_dir = "{tmp_path / 'out'}"
[texture.s]
path = "{tmp_path / 's.png'}"
channels = 3
build_mipmaps = true
colorspace = "sRGB"
[[task]]
type = "convert_texture"
task_dir = "conv"
textures = ["texture.s"]
"""
    cfg = tmp_path / "conv.toml"
    cfg.write_text(text)
    r = subprocess.run([str(TOOL), "-c", str(cfg)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    w, h, c, dt, levels, data = parse_serialized(tmp_path / "out" / "conv" / "s.bin")
    assert (w, h, c, dt, levels) == (21, 37, 3, 0, 7)
    np.testing.assert_array_equal(np.frombuffer(data[: 37 * 21 * 3], np.uint8).reshape(37, 21, 3), img)
    # the levels are the reference's chain: the oracle's, read through the sRGB table
    osc = ora.OracleScene(np.float32([[0, 0, -9], [1, 0, -9], [0, 1, -9]]), np.uint32([[0, 1, 2]]), np.uint32([0]),
                          np.uint32([0]), np.zeros(1, api.MATERIAL_DTYPE), np.zeros(0, api.LIGHT_DTYPE),
                          textures=[dict(data=img, srgb=True, mipmaps=True)])
    srgb, _ = ora.tex_tables()
    off, lw, lh = 0, 21, 37
    for lvl in range(levels):
        n = lw * lh * 3
        mine = srgb[np.frombuffer(data[off:off + n], np.uint8)].reshape(lh, lw, 3)
        np.testing.assert_array_equal(mine, osc.texture_level(0, lvl), err_msg=f"level {lvl}")
        off += n
        lw, lh = max(1, (lw + 1) // 2), max(1, (lh + 1) // 2)
    # read back: serialized textures load as linear, every level kept
    r, d = dump(tmp_path, config(tmp_path, {"t": f'path = "{tmp_path / "out" / "conv" / "s.bin"}"\nserialized = true'},
                                 bsdf_extra=BASECOLOR_TEX))
    assert r.returncode == 0, r.stderr
    t0 = read_dump(d)["textures"][0]
    assert (t0["w"], t0["h"], t0["c"], t0["cs"], t0["levels"]) == (21, 37, 3, api.KS_COLORSPACE_LINEAR, 7)
    assert t0["bytes"] == data


# ---- shader fields, normal maps, texcoords ------------------------------------------------------------

def test_field_tables_and_normal_map(tmp_path):
    img = np.random.default_rng(1).integers(0, 256, (4, 4, 4)).astype(np.uint8)
    imf.write_png_any(tmp_path / "t.png", img, 6, 8)
    bsdf = ('basecolor = { type = "texture", map = "texture.t", sampler = { type = "ewa", anisotropy = 4.0, '
            'wrap_mode_u = "clamp" }, flip_v = false, swizzle = [2, 1, 0], uv_scale = [2, 3], uv_offset = [0.5, 0.25], '
            'scale = [1, 0.5, 0.25] }\n'
            'roughness = { type = "texture", map = "texture.t", sampler = { type = "cubic", kernel = "catmull_rom" }, '
            'swizzle = 3 }')
    extra = ('[normal_map.n]\nmap = { type = "texture", map = "texture.t", sampler = { type = "nearest" } }\n'
             'space = "tangent"\nstrength = 0.5\nrange01 = false\n')
    text = config(tmp_path, {"t": f'path = "{tmp_path / "t.png"}"\nchannels = 4\nbuild_mipmaps = true\ncolorspace = "sRGB"'},
                  bsdf_extra="", material_extra='normal_map = "normal_map.n"', extra_assets=extra)
    text = text.replace('roughness = { type = "constant", value = 0.5 }\n', bsdf + "\n")
    r, d = dump(tmp_path, text)
    assert r.returncode == 0, r.stderr
    res = read_dump(d)
    assert res["field_of"].tolist() == [0] and len(res["textures"]) == 1  # one texture, loaded once
    f = res["fields"][0]
    bc = f.basecolor
    assert (bc.texture, bc.sampler, bc.wrap_u, bc.wrap_v, bc.flip_v) == (0, api.KS_SAMPLER_EWA, api.KS_WRAP_CLAMP,
                                                                       api.KS_WRAP_REPEAT, 0)
    assert bc.anisotropy == 4.0 and list(bc.swizzle)[:3] == [2, 1, 0]
    assert list(bc.uv_scale) == [2, 3] and list(bc.uv_offset) == [0.5, 0.25] and list(bc.scale)[:3] == [1, 0.5, 0.25]
    ro = f.roughness
    assert (ro.sampler, ro.cubic_kernel, ro.swizzle[0], ro.flip_v) == (api.KS_SAMPLER_CUBIC, api.KS_CUBIC_CATMULL_ROM, 3, 1)
    assert f.metallic.texture == -1 and f.ior.value[0] == 1.5
    assert f.has_normal_map == 1 and f.normal_map.sampler == api.KS_SAMPLER_NEAREST
    assert f.normal_space == api.KS_NORMAL_MAP_TANGENT and f.normal_strength == 0.5 and f.normal_range01 == 0
    # the OBJ's texture coordinates reach the scene, per remapped vertex (uv = xy on this quad)
    verts = np.fromfile(d / "vertices.f32", np.float32).reshape(-1, 3)
    np.testing.assert_array_equal(res["texcoords"], verts[:, :2])
    assert res["textures"][0]["cs"] == api.KS_COLORSPACE_SRGB and res["textures"][0]["mips"] == 1


@pytest.mark.parametrize("bsdf,msg", [
    ('basecolor = { type = "texture", map = "texture.t", sampler = { type = "bicubic" } }', "unknown texture sampler"),
    ('basecolor = { type = "texture", map = "texture.t", sampler = { type = "linear" }, swizzle = [0, 1, 5] }', "swizzle"),
    ('basecolor = { type = "texture", map = "texture.nope", sampler = { type = "linear" } }', "texture.nope"),
    ('basecolor = { type = "procedural" }', "unknown shader field type"),
])
def test_field_errors(tmp_path, bsdf, msg):
    imf.write_png_any(tmp_path / "t.png", np.zeros((2, 2, 3), np.uint8), 2, 8)
    r, _ = dump(tmp_path, config(tmp_path, {"t": f'path = "{tmp_path / "t.png"}"\nchannels = 3\nbuild_mipmaps = false'},
                                 bsdf_extra=bsdf))
    assert r.returncode != 0 and msg in r.stderr, r.stderr


def test_lz4_roundtrip_through_the_tool(tmp_path):
    """Serialized f32 / u16 textures with the chain written from level 0 only (build_mipmaps off)."""
    rng = np.random.default_rng(2)
    f = rng.uniform(0, 1, (9, 7, 4)).astype(np.float32)
    imf.write_exr(tmp_path / "f.exr", f, channels="RGBA")
    text = f"""
output_dir = "{tmp_path / 'out'}"
[texture.f]
path = "{tmp_path / 'f.exr'}"
channels = 2
build_mipmaps = false
[[task]]
type = "convert_texture"
task_dir = "conv"
textures = ["texture.f"]
"""
    (tmp_path / "c.toml").write_text(text)
    r = subprocess.run([str(TOOL), "-c", str(tmp_path / "c.toml")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    w, h, c, dt, levels, data = parse_serialized(tmp_path / "out" / "conv" / "f.bin")
    assert (w, h, c, dt, levels) == (7, 9, 2, 2, 1)
    np.testing.assert_array_equal(np.frombuffer(data, np.float32).reshape(9, 7, 2), f[..., :2])
