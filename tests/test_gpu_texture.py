"""Textured scenes on the GPU against the oracle (texture.cpp, shader_field.h, normal_map.cpp,
opacity_map.cpp, light.cpp:307-523): the textured box exercises every sampler (nearest, trilinear,
Catmull-Rom / B-spline cubic, EWA) over u8 / u16 / f32 textures in linear and sRGB space, a normal map,
cut-out and constant opacity maps in both traversal kernels, textured emission in the light table, in
emitter hits and in light samples, uv-less geometry and camera-ray uv partials. Bars: the light table
within 1e-5, per-sample radiance within 1e-3 on > 97% of samples (the same Sobol points drive both;
paths part only where float rounding flips a decision), images within relMSE 1e-3 at 16 spp.

The stochastic opacity test hashes the bits of the ray origin and direction (opacity_map.cpp:20-31), and
the GPU's camera and bounce rays differ from the oracle's in the last bit, so the per-sample and image
comparisons use the box with that quad opaque; the stochastic test itself is checked on identical rays
(decisions equal) and for its transmitted fraction."""

import numpy as np
import pytest

from conftest import rel_mse
from texture_scenes import textured_box

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def box():
    return textured_box(stochastic_opacity=False)


@pytest.fixture(scope="module")
def box_stochastic():
    return textured_box()


def test_textured_light_table(device, ks, ora, box):
    sc = ks.Scene(device, box)
    gp, gpow, gcdf = sc.light_table()
    op, opow, ocdf = ora.OracleScene.from_host_scene(box).light_table()
    _log("lights", gp, gpow, opow)
    assert np.array_equal(gp, op)
    np.testing.assert_allclose(gpow, opow, rtol=1e-5)
    np.testing.assert_allclose(gcdf, ocdf, rtol=1e-5, atol=1e-7)
    sc.close()


@pytest.mark.parametrize("scale_ray_diff", [True, False])
def test_textured_trace_samples(device, ks, ora, box, scale_ray_diff):
    sc = ks.Scene(device, box)
    osc = ora.OracleScene.from_host_scene(box)
    W = H = 192
    spp = 16
    rng = np.random.default_rng(3)
    n = 20000
    px = rng.integers(0, W, n).astype(np.int32)
    py = rng.integers(0, H, n).astype(np.int32)
    s = rng.integers(0, spp, n).astype(np.int32)
    kw = dict(bounces=6, render_width=W, render_height=H, spp=spp, scale_ray_diff=scale_ray_diff)
    g = sc.trace_samples(box.camera, ks.small_pt_params(**kw), px, py, s)
    o = osc.trace_samples(box.camera, ora.params(**kw), px, py, s)
    close = np.all(np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o)), axis=1)
    _log("textured per-sample agreement", close.mean(), "mean", g[:, :3].mean(axis=0), o[:, :3].mean(axis=0))
    assert close.mean() > 0.97
    assert np.array_equal(g[:, 3], o[:, 3])
    sc.close()


def test_textured_render_matches_oracle(device, ks, ora, box):
    W = H = 128
    kw = dict(bounces=6, render_width=W, render_height=H, spp=16, rng_seed=0)
    sc = ks.Scene(device, box)
    film = ks.Film(device, W, H)
    st = ks.small_pt_render(sc, box.camera, ks.small_pt_params(**kw), film)
    grgb, gw = film.download()
    orgb, ow = ora.OracleScene.from_host_scene(box).render(box.camera, ora.params(**kw))
    assert np.array_equal(gw, ow)
    gimg = grgb / np.maximum(gw[..., :1], 1)
    oimg = orgb / np.maximum(ow[..., :1], 1)
    e = rel_mse(gimg, oimg)
    _log("textured box relMSE gpu vs oracle", e, "mean", gimg.mean(axis=(0, 1)), oimg.mean(axis=(0, 1)),
         "nonfinite", st["nonfinite"])
    assert e < 1e-3
    assert st["nonfinite"] == 0
    film.close()
    sc.close()


def test_opacity_queries(device, ks, ora, box_stochastic):
    """ks_intersect / ks_occluded apply the opacity filter as the oracle's closest / occluded do, and the
    half-opaque quad stops half of the rays that reach it."""
    box = box_stochastic
    sc = ks.Scene(device, box)
    osc = ora.OracleScene.from_host_scene(box)
    rng = np.random.default_rng(4)
    n = 50000
    o = np.tile(np.float32([0.35, 0.3, 1.5]), (n, 1)) + rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    tgt = np.stack([rng.uniform(0.1, 0.95, n), rng.uniform(0.05, 0.6, n), np.full(n, 0.3)], -1).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = o, d, 100.0
    g = sc.intersect(rays)
    oo = osc.intersect(rays)
    gp, op = g[:, 3].view(np.int32), oo[:, 3].view(np.int32)
    agree = (gp == op).mean()
    _log("alpha-tested closest-hit agreement", agree, "cut-out / half quads hit", np.isin(gp, [12, 13, 14, 15]).mean())
    assert agree > 0.999
    occ = sc.occluded(rays)
    assert (occ.astype(bool) == (op >= 0)).mean() > 0.999
    # rays aimed at the middle of the constant-0.5 quad (triangles 14, 15) from spread-out origins
    m = 40000
    o2 = np.stack([rng.uniform(0.6, 0.8, m), rng.uniform(0.2, 0.35, m), np.full(m, 1.5)], -1).astype(np.float32)
    t2 = np.stack([rng.uniform(0.6, 0.8, m), rng.uniform(0.2, 0.35, m), np.full(m, 0.5)], -1).astype(np.float32)
    d2 = t2 - o2
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    r2 = np.zeros((m, 8), np.float32)
    r2[:, 0:3], r2[:, 4:7], r2[:, 7] = o2, d2, 100.0
    g2 = sc.intersect(r2)[:, 3].view(np.int32)
    stopped = np.isin(g2, [14, 15]).mean()
    _log("fraction stopped by the 0.5-opacity quad", stopped)
    assert abs(stopped - 0.5) < 4 * np.sqrt(0.25 / m)
    assert np.array_equal(g2, osc.intersect(r2)[:, 3].view(np.int32))
    sc.close()


def scene_from_dump(ks, d, camera):
    """The HostScene that `ks_render --dump-scene` wrote to directory d."""
    import ctypes as C

    rd = lambda n, dt: np.fromfile(d / n, dt)
    v = rd("vertices.f32", np.float32).reshape(-1, 3)
    nrm = rd("normals.f32", np.float32)
    mats = np.frombuffer((d / "materials.bin").read_bytes(), ks.MATERIAL_DTYPE).copy()
    lights = np.frombuffer((d / "lights.bin").read_bytes(), ks.LIGHT_DTYPE).copy()
    field_of = rd("field_of.i32", np.int32)
    raw = (d / "fields.bin").read_bytes()
    sz = C.sizeof(ks.MaterialFields)
    fields = {m: ks.MaterialFields.from_buffer_copy(raw, int(k) * sz) for m, k in enumerate(field_of) if k >= 0}
    textures = []
    dtypes = {0: np.uint8, 1: np.uint16, 2: np.float32}
    for i, line in enumerate(l for l in (d / "textures.txt").read_text().split("\n") if l.strip()):
        w, h, c, dt, cs, levels, mips = map(int, line.split())
        data = np.fromfile(d / f"texture_{i}.bin", dtypes[dt])
        lv, off = [], 0
        for _ in range(levels):
            lv.append(data[off:off + w * h * c].reshape(h, w, c))
            off += w * h * c
            w, h = max(1, (w + 1) // 2), max(1, (h + 1) // 2)
        textures.append(dict(data=lv if levels > 1 else lv[0], srgb=cs == 1, mipmaps=bool(mips)))
    tc = rd("texcoords.f32", np.float32)
    gf = rd("geom_flags.u8", np.uint8)
    return ks.HostScene(vertices=v, indices=rd("indices.u32", np.uint32).reshape(-1, 3),
                        tri_geom=rd("tri_geom.u32", np.uint32), geom_material=rd("geom_material.u32", np.uint32),
                        materials=mats, lights=lights, camera=camera, normals=nrm.reshape(-1, 3) if nrm.size else None,
                        geom_flags=gf if gf.size else None, texcoords=tc.reshape(-1, 2) if tc.size else None,
                        textures=textures or None, material_fields=fields or None)


def test_ks_render_textured_toml_scene(tmp_path, device, ks):
    """A TOML scene with PNG / EXR textures, a normal map and a textured emitter: ks_render's EXR equals
    the API render of the scene it resolved."""
    import subprocess
    from pathlib import Path

    import image_files as imf
    from image_files import read_exr_rgb

    tool = Path(__file__).resolve().parents[1] / "keepsake_amd" / "_lib" / "ks_render"
    rng = np.random.default_rng(0)
    imf.write_png_any(tmp_path / "floor.png", rng.integers(0, 256, (40, 40, 3)), 2, 8)
    yy, xx = np.mgrid[0:32, 0:32]
    nm = np.stack([0.3 * np.sin(xx / 5.0), 0.3 * np.cos(yy / 4.0), np.ones_like(xx, np.float64)], -1)
    nm /= np.linalg.norm(nm, axis=-1, keepdims=True)
    imf.write_png_any(tmp_path / "nrm.png", np.round((nm * 0.5 + 0.5) * 255).astype(np.int64), 2, 8)
    imf.write_exr(tmp_path / "em.exr", np.stack([6 + 4 * np.sin(xx / 3.0), 5 + 0 * xx, 3 + 2 * (yy > 16)], -1).astype(np.float32))
    (tmp_path / "floor.obj").write_text("v 0 0 0\nv 0 0 1\nv 1 0 1\nv 1 0 0\nvt 0 0\nvt 0 2\nvt 2 2\nvt 2 0\n"
                                        "f 1/1 2/2 3/3 4/4\n"
                                        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                                        "o wall\nf 5/5 6/6 7/7 8/8\n"
                                        "v 0.3 0.99 0.3\nv 0.7 0.99 0.3\nv 0.7 0.99 0.7\nv 0.3 0.99 0.7\n"
                                        "o light\nf 9/5 10/6 11/7 12/8\n")
    const = lambda v: f'{{ type = "constant", value = {v} }}'
    bsdf = lambda name, bc: (f'[bsdf.{name}]\ntype = "principled_bsdf"\nbasecolor = {bc}\nroughness = {const(0.6)}\n'
                             f'metallic = {const(0.0)}\nior = {const(1.5)}\nspecular_r0_mul = {const(0.5)}\n'
                             f'specular_trans = {const(0.0)}\ndiffuse_trans = {const(0.0)}\ndiffuse_trans_fwd = {const(0.0)}\n')
    cfg = tmp_path / "c.toml"
    cfg.write_text(f"""
output_dir = "{tmp_path / 'out'}"
[texture.floor]
path = "{tmp_path / 'floor.png'}"
channels = 3
build_mipmaps = true
colorspace = "sRGB"
[texture.nrm]
path = "{tmp_path / 'nrm.png'}"
channels = 3
build_mipmaps = false
[texture.em]
path = "{tmp_path / 'em.exr'}"
channels = 3
build_mipmaps = true
[normal_map.wall]
map = {{ type = "texture", map = "texture.nrm", sampler = {{ type = "linear" }} }}
strength = 0.9
[mesh_asset.room]
path = "{tmp_path / 'floor.obj'}"
format = "obj"
load_materials = false
twosided = true
{bsdf("floor", '{ type = "texture", map = "texture.floor", sampler = { type = "ewa" } }')}
{bsdf("wall", const([0.7, 0.7, 0.7]))}
{bsdf("dark", const([0.0, 0.0, 0.0]))}
[material.floor]
bsdf = "bsdf.floor"
[material.wall]
bsdf = "bsdf.wall"
normal_map = "normal_map.wall"
[material.light]
bsdf = "bsdf.dark"
emission = {{ type = "texture", map = "texture.em", sampler = {{ type = "cubic", kernel = "catmull_rom" }} }}
[[task]]
type = "small_pt"
task_dir = "tex"
object = "mesh_asset.room"
material = [["material.floor", "material.wall", "material.light"]]
sky = {{ ambient = [0.1, 0.1, 0.12] }}
bounces = 4
render_width = 64
render_height = 64
spp = 8
[task.camera]
type = "perspective"
pos = [0.5, 0.5, 2.0]
target = [0.5, 0.3, 0.0]
up = [0.0, 1.0, 0.0]
vfov = 45.0
aspect = 1.0
""")
    r = subprocess.run([str(tool), "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    _log(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    exr, _ = read_exr_rgb(tmp_path / "out" / "tex" / "small_pt_spp8.exr")
    r = subprocess.run([str(tool), "-c", str(cfg), "--dump-scene", str(tmp_path / "dump")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cam = dict(pos=(0.5, 0.5, 2.0), target=(0.5, 0.3, 0.0), up=(0.0, 1.0, 0.0), vfov=45.0, aspect=1.0)
    hs = scene_from_dump(ks, tmp_path / "dump" / "task_0", cam)
    assert hs.material_fields is not None and len(hs.textures) == 3
    sc = ks.Scene(device, hs)
    film = ks.Film(device, 64, 64)
    st = ks.small_pt_render(sc, cam, ks.small_pt_params(bounces=4, render_width=64, render_height=64, spp=8), film)
    ref = film.composite()
    _log("ks_render vs api on the textured TOML scene: max abs diff", float(np.abs(exr - ref).max()),
         "mean", ref.mean(axis=(0, 1)), "nonfinite", st["nonfinite"])
    assert np.array_equal(exr, ref) and ref.mean() > 0.01
    film.close()
    sc.close()
