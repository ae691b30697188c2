"""Mesh-asset scenes on the GPU.

1. The procedural Cornell box written out as an OBJ with one shape per geometry, loaded back through
   the TOML path (object + material list + [material.*]/[bsdf.*] tables, mesh-light emission), renders
   the same EXR bit for bit as the procedural scene: loading changes the vertex numbering, not the
   triangles, so the BVH, the light table and every sample agree.
2. Per-geometry flags (one-sided shading frames, geometric instead of vertex normals, geometry.cpp:68-163)
   match the CPU oracle at equal spp, and visibly change the image, so the flags are not ignored."""

import copy
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_mse
from image_files import read_exr_rgb

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "keepsake_amd" / "_lib" / "ks_render"


def read_image(path):
    return read_exr_rgb(path)[0]


def _f(x):
    return repr(float(x))


def _vec(v):
    return "[" + ", ".join(_f(x) for x in v) + "]"


def write_obj(hs, path):
    lines = [f"v {_f(x)} {_f(y)} {_f(z)}" for x, y, z in hs.vertices]
    for g in range(hs.geom_material.shape[0]):
        lines.append(f"o geom{g}")
        for t in np.nonzero(hs.tri_geom == g)[0]:
            a, b, c = (int(i) + 1 for i in hs.indices[t])
            lines.append(f"f {a} {b} {c}")
    Path(path).write_text("\n".join(lines) + "\n")


def material_tables(hs):
    out = []
    for i, m in enumerate(hs.materials):
        def one(k):
            return f'{{ type = "constant", value = {_f(m[k])} }}'

        out.append(f"""
[bsdf.m{i}]
type = "principled_bsdf"
basecolor = {{ type = "constant", value = {_vec(m["basecolor"])} }}
roughness = {one("roughness")}
metallic = {one("metallic")}
ior = {one("ior")}
specular_r0_mul = {one("specular_r0_mul")}
specular_trans = {one("specular_trans")}
diffuse_trans = {one("diffuse_trans")}
diffuse_trans_fwd = {one("diffuse_trans_fwd")}
emissive = {{ type = "constant", value = [0.0, 0.0, 0.0] }}
[material.m{i}]
bsdf = "bsdf.m{i}"
""")
        if m["has_emission"]:
            out.append(f"emission = {_vec(m['emission'])}\n")
    return "".join(out)


def test_obj_cornell_renders_like_procedural(tmp_path, ks, cornell):
    assert len(cornell.lights) == 0 and cornell.normals is None
    write_obj(cornell, tmp_path / "cornell.obj")
    cam = cornell.camera
    camera = (f'camera = {{ type = "perspective", pos = {_vec(cam["pos"])}, target = {_vec(cam["target"])}, '
              f'up = {_vec(cam["up"])}, vfov = {_f(cam["vfov"])}, aspect = {_f(cam["aspect"])} }}')
    common = f"""bounces = 8
render_width = 96
render_height = 64
spp = 16
spp_prog_interval = 16
{camera}
"""
    mats = ", ".join(f'"material.m{int(m)}"' for m in cornell.geom_material)
    cfg = tmp_path / "c.toml"
    cfg.write_text(f"""output_dir = "{tmp_path / 'out'}"
[mesh_asset.box]
path = "{tmp_path / 'cornell.obj'}"
load_materials = false
twosided = true
{material_tables(cornell)}
[[task]]
type = "small_pt"
task_dir = "procedural"
procedural = "cornell"
{common}
[[task]]
type = "small_pt"
task_dir = "obj"
object = "mesh_asset.box"
material = [[{mats}]]
{common}
""")
    print("ks_render on", cfg, flush=True)
    r = subprocess.run([str(TOOL), "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    a = read_image(tmp_path / "out" / "procedural" / "small_pt_spp16.exr")
    b = read_image(tmp_path / "out" / "obj" / "small_pt_spp16.exr")
    assert a.mean() > 0.01  # the mesh light lit the box
    assert np.array_equal(a, b)


def _flagged_cornell(ks, cornell):
    hs = copy.deepcopy(cornell)
    rng = np.random.default_rng(7)
    # vertex normals: the face normal of each triangle's plane, tilted at random by up to ~17 degrees
    v = hs.vertices.astype(np.float64)
    nrm = np.zeros_like(v)
    for t in hs.indices:
        a, b, c = v[t]
        n = np.cross(b - a, c - a)
        nrm[t] += n / np.linalg.norm(n)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm += 0.3 * rng.normal(size=nrm.shape)
    hs.normals = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    G = hs.geom_material.shape[0]
    flags = np.zeros(G, np.uint8)
    flags[0::3] |= ks.KS_GEOM_ONE_SIDED
    flags[1::2] |= ks.KS_GEOM_FLAT
    # reverse the winding of one one-sided geometry so its front faces point away from the room
    g = 0
    sel = hs.tri_geom == g
    idx = hs.indices.copy()
    idx[sel] = idx[sel][:, [0, 2, 1]]
    hs.indices = idx
    hs.geom_flags = flags
    return hs


def _render(ks, ora, device, hs, kw):
    sc = ks.Scene(device, hs)
    film = ks.Film(device, kw["render_width"], kw["render_height"])
    st = ks.small_pt_render(sc, hs.camera, ks.small_pt_params(**kw), film)
    grgb, gw = film.download()
    film.close()
    sc.close()
    orgb, ow = ora.OracleScene.from_host_scene(hs).render(hs.camera, ora.params(**kw))
    return grgb, gw, orgb, ow, st


def test_geometry_flags_match_oracle(device, ks, ora, cornell):
    kw = dict(bounces=8, render_width=128, render_height=128, spp=16, rng_seed=0)
    hs = _flagged_cornell(ks, cornell)
    print("flags", hs.geom_flags.tolist(), flush=True)
    grgb, gw, orgb, ow, st = _render(ks, ora, device, hs, kw)
    assert np.array_equal(gw, ow)
    gimg = grgb / np.maximum(gw[..., :1], 1)
    oimg = orgb / np.maximum(ow[..., :1], 1)
    e = rel_mse(gimg, oimg)
    print("flagged cornell relMSE gpu vs oracle", e, "nonfinite", st["nonfinite"], flush=True)
    assert e < 1e-3 and st["nonfinite"] == 0
    # the same scene without flags renders differently on both sides
    plain = copy.deepcopy(hs)
    plain.geom_flags = None
    prgb, pw, porgb, pow_, _ = _render(ks, ora, device, plain, kw)
    pimg = prgb / np.maximum(pw[..., :1], 1)
    d = rel_mse(pimg, gimg)
    print("flags vs no flags relMSE", d, flush=True)
    assert d > 20 * max(e, 1e-5)


def test_gltf_cornell_renders_like_procedural(tmp_path, ks, cornell):
    """The Cornell box as one glTF mesh (a primitive per geometry, glTF materials with factors only, the
    light's radiance as emissiveFactor) renders the procedural scene's image bit for bit."""
    from test_gltf_asset import cornell_gltf

    cornell_gltf(cornell).write_glb(tmp_path / "cornell.glb")
    cam = cornell.camera
    camera = (f'camera = {{ type = "perspective", pos = {_vec(cam["pos"])}, target = {_vec(cam["target"])}, '
              f'up = {_vec(cam["up"])}, vfov = {_f(cam["vfov"])}, aspect = {_f(cam["aspect"])} }}')
    common = f"bounces = 8\nrender_width = 96\nrender_height = 64\nspp = 16\nspp_prog_interval = 16\n{camera}\n"
    cfg = tmp_path / "c.toml"
    cfg.write_text(f"""output_dir = "{tmp_path / 'out'}"
[compound_mesh_asset.box]
path = "{tmp_path / 'cornell.glb'}"
load_materials = true
[[task]]
type = "small_pt"
task_dir = "procedural"
procedural = "cornell"
{common}
[[task]]
type = "small_pt"
task_dir = "gltf"
compound_object = "compound_mesh_asset.box"
{common}
""")
    print("ks_render on", cfg, flush=True)
    r = subprocess.run([str(TOOL), "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    a = read_image(tmp_path / "out" / "procedural" / "small_pt_spp16.exr")
    b = read_image(tmp_path / "out" / "gltf" / "small_pt_spp16.exr")
    assert a.mean() > 0.01
    assert np.array_equal(a, b)
