"""The C++ host side (ks_render, keepsake_amd/host): TOML parsing and the small_pt task's key handling,
checked through `ks_render --print-config` (no GPU needed). Defaults follow small_pt.cpp:227-241."""

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "keepsake_amd" / "_lib" / "ks_render"

BASE = """
output_dir = "{out}"
[[task]]
type = "small_pt"
procedural = "cornell"
bounces = 8
render_width = 64
render_height = 32
spp = 4
"""


def resolve(tmp_path, text):
    cfg = tmp_path / "c.toml"
    cfg.write_text(text.replace("{out}", str(tmp_path / "out")))
    r = subprocess.run([str(TOOL), "-c", str(cfg), "--print-config"], capture_output=True, text=True)
    return r


def tasks(tmp_path, text):
    r = resolve(tmp_path, text)
    assert r.returncode == 0, r.stderr
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def test_defaults_match_small_pt(tmp_path):
    (t,) = tasks(tmp_path, BASE)
    assert t["scene"] == "cornell" and t["n_tris"] == 36
    assert t["clamp_direct"] == 0 and t["clamp_indirect"] == 10  # small_pt.cpp:230-231
    assert t["spp_prog_interval"] == 32 and t["rng_seed"] == 0
    assert t["include_background"] is True and t["crop"] == [0, 0, 0, 0]
    assert t["mis_mode"] == 0
    assert t["bvh_builder"] == 0  # the host SBVH
    # without a [task.camera] table the procedural scene's own camera is used
    assert t["camera"]["vfov_deg"] == pytest.approx(40.0)


def test_bvh_builder_key(tmp_path):
    (t,) = tasks(tmp_path, BASE + 'bvh_builder = "device"\n')
    assert t["bvh_builder"] == 1


def test_backdrop_disables_background(tmp_path):
    (t,) = tasks(tmp_path, BASE + "backdrop = [0.1, 0.2, 0.3]\n")
    assert t["include_background"] is False
    assert t["backdrop"] == pytest.approx([0.1, 0.2, 0.3])


def test_override_merges_tables(tmp_path):
    text = BASE + """
[[task]]
override = 0
spp = 16
crop_start_x = 8
crop_width = 16
crop_height = 8
"""
    a, b = tasks(tmp_path, text)
    assert a["spp"] == 4 and b["spp"] == 16
    assert b["bounces"] == 8 and b["crop"] == [8, 0, 16, 8]


def test_camera_keys(tmp_path):
    text = BASE + """
[task.camera]
type = "perspective"
pos = [1, 2, 3]
target = [0, 0, 0]
up = [0, 2, 0]   # normalised on load
hfov = 60.0
aspect = 2.0
"""
    (t,) = tasks(tmp_path, text)
    c = t["camera"]
    assert c["pos"] == [1, 2, 3] and c["up"] == [0, 1, 0]
    # camera.cpp:196-197: vfov = 2 atan(tan(hfov/2) / aspect)
    assert c["vfov_deg"] == pytest.approx(np.rad2deg(2 * np.arctan(np.tan(np.deg2rad(30.0)) / 2.0)), rel=1e-5)


def test_to_world_transform(tmp_path):
    text = BASE + """
[task.camera]
type = "perspective"
vfov = 45
aspect = 1.0
to_world = { rotation = { euler = [0, 90, 0] }, translation = [1, 2, 3] }
"""
    (t,) = tasks(tmp_path, text)
    m = np.array(t["camera"]["to_world"]).reshape(4, 4)
    ry = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]])  # rotation of +90 degrees about y
    assert np.allclose(m[:3, :3], ry, atol=1e-6)
    assert np.allclose(m[:3, 3], [1, 2, 3]) and m[3, 3] == 1
    text2 = BASE + """
[task.camera]
type = "perspective"
vfov = 45
aspect = 1.0
to_world.matrix_cm = [1,0,0,0, 0,1,0,0, 0,0,1,0, 5,6,7,1]
"""
    (t2,) = tasks(tmp_path, text2)
    m2 = np.array(t2["camera"]["to_world"]).reshape(4, 4)
    assert np.allclose(m2[:3, 3], [5, 6, 7])


def test_lights(tmp_path):
    text = BASE + """
sky = { ambient = [0.5, 0.5, 0.5] }
light = [
  { type = "directional", L = [1, 1, 1], dir = [0, -2, 0] },  # trailing comma and comments
  { type = "point", color = [1, 1, 1], power = 12.566370614359172, pos = [0, 1, 0] },
]
"""
    (t,) = tasks(tmp_path, text)
    assert t["n_lights"] == 3


def test_toml_syntax(tmp_path):
    text = """
# comment
output_dir = '{out}'
[[task]]
"type" = "small_pt"   # quoted key
procedural = "spheres_small"
bounces = 1_2
render_width = 1_920
render_height = 1080
spp = 64
clamp_indirect = 2.5e1
scale_ray_diff = false
"""
    (t,) = tasks(tmp_path, text)
    assert t["bounces"] == 12 and t["render_width"] == 1920 and t["clamp_indirect"] == 25.0


@pytest.mark.parametrize(
    "extra,msg",
    [
        ('object = "mesh_asset.teapot"\n', "unknown asset 'mesh_asset.teapot'"),
        ('compound_object = "compound_mesh_asset.x"\n', "unknown asset 'compound_mesh_asset.x'"),
        ('camera_animation = "anim"\n', "camera_animation"),
        ('pixel_filter = { type = "mitchell" }\n', "Invalid pixel filter type"),
        ('mis_mode = "fancy"\n', "mis_mode"),
        ('bvh_builder = "gpu"\n', "bvh_builder"),
        ('spp = "many"\n', "duplicate key"),
    ],
)
def test_errors_are_reported(tmp_path, extra, msg):
    r = resolve(tmp_path, BASE + extra)
    assert r.returncode != 0
    assert msg in r.stderr


def test_missing_required_key(tmp_path):
    r = resolve(tmp_path, BASE.replace("spp = 4\n", ""))
    assert r.returncode != 0 and "spp" in r.stderr


def test_guide_table(tmp_path):
    text = BASE + """mis_mode = "consistent"
[task.guide]
fraction = 0.25
train_spp = 32
learning_rate = 5e-3
log2_table_size = 18
max_bounce = 2
"""
    (t,) = tasks(tmp_path, text)
    assert t["guided"] is True
    g = t["guide"]
    assert g["max_bounce"] == 2
    assert g["fraction"] == pytest.approx(0.25) and g["train_spp"] == 32
    assert g["learning_rate"] == pytest.approx(5e-3) and g["log2_table_size"] == 18
    (plain,) = tasks(tmp_path, BASE)
    assert plain["guided"] is False


def test_guide_fraction_zero_is_kept(tmp_path):
    """fraction = 0 (guide trained, never sampled) stays 0; a [task.guide] table without the key gets the
    library default 0.5."""
    text = BASE + 'mis_mode = "consistent"\n[task.guide]\n'
    (t,) = tasks(tmp_path, text + "fraction = 0.0\n")
    assert t["guide"]["fraction"] == 0.0
    (t,) = tasks(tmp_path, text + "train_spp = 4\n")
    assert t["guide"]["fraction"] == pytest.approx(0.5)


def test_guide_needs_consistent_mis(tmp_path):
    r = resolve(tmp_path, BASE + "[task.guide]\nfraction = 0.5\n")
    assert r.returncode != 0 and "consistent" in r.stderr


def test_guide_inference_precision(tmp_path):
    text = BASE + """mis_mode = "consistent"
[task.guide]
inference = "fp8"
"""
    (t,) = tasks(tmp_path, text)
    assert t["guide"]["inference_fp8"] == 1
    r = resolve(tmp_path, text.replace('"fp8"', '"fp4"'))
    assert r.returncode != 0 and "bf16" in r.stderr


def test_sky_map(tmp_path):
    import image_files

    rgb = np.random.default_rng(0).uniform(0.1, 4.0, (6, 10, 3)).astype(np.float32)
    image_files.write_hdr(tmp_path / "sky.hdr", image_files.to_rgbe(rgb))
    image_files.write_exr(tmp_path / "sky.exr", rgb)
    text = BASE + f'sky = {{ map = "{tmp_path}/sky.hdr", strength = 2.0, transform_y_up = false, to_world = {{ rotation = {{ euler = [0, 90, 0] }} }} }}\n'
    (t,) = tasks(tmp_path, text)
    assert t["n_lights"] == 1 and t["sky_map"] == [10, 6, 2, 0]
    text = BASE + f'light = [ {{ type = "sky", map = "{tmp_path}/sky.exr" }} ]\n'
    (t,) = tasks(tmp_path, text)
    assert t["n_lights"] == 1 and t["sky_map"] == [10, 6, 1, 1]  # defaults: strength 1, y up
    r = resolve(tmp_path, BASE + f'sky = {{ map = "{tmp_path}/sky.png" }}\n')
    assert r.returncode != 0 and "only .hdr or .exr" in r.stderr
    r = resolve(tmp_path, BASE + f'sky = {{ map = "{tmp_path}/sky.hdr" }}\n'
                f'light = [ {{ type = "sky", map = "{tmp_path}/sky.exr" }} ]\n')
    assert r.returncode != 0 and "at most one sky map" in r.stderr


@pytest.mark.parametrize("text,want", [
    ('pixel_filter = { type = "gaussian" }\n', (1, 3.0, 0.5)),
    ('pixel_filter = { type = "gaussian", width = 2.0, sigma = 0.3 }\n', (1, 2.0, 0.3)),
    ('pixel_filter = { type = "blackman_harris" }\n', (2, 3.0, 0.0)),
    ('pixel_filter = { type = "box", width = 2.0 }\n', (0, 2.0, 0.0)),
    ("", (0, 0.0, 0.0)),
])
def test_pixel_filter_key(tmp_path, text, want):
    """create_pixel_filter's keys and defaults (render_target.cpp:28-44)."""
    (t,) = tasks(tmp_path, BASE + text)
    f = t["pixel_filter"]
    assert (f["type"], f["width"], f["sigma"]) == pytest.approx(want)


def test_pixel_filter_unknown_type(tmp_path):
    r = resolve(tmp_path, BASE + 'pixel_filter = { type = "mitchell" }\n')
    assert r.returncode != 0 and "Invalid pixel filter type" in r.stderr
