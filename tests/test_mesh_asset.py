"""Mesh assets on the TOML path (keepsake_amd/host/mesh_asset.cpp), checked without a GPU through
`ks_render --dump-scene`: OBJ shapes, (v, vt, vn) remapping in first-use order and triangulation
(mesh_asset.cpp:70-177), the binary mesh asset layout (mesh_asset.cpp:180-241) read back by an
independent Python reader, `to_world` (MeshData::transform, geometry.cpp:15-32), material tables
(create_material material.cpp:169-177, create_principled_bsdf principled_bsdf.cpp:707-732) and the
material list (assign_material_list mesh_asset.cpp:285-299)."""

import json
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from keepsake_amd.api import KS_GEOM_FLAT, KS_GEOM_ONE_SIDED, LIGHT_DTYPE, MATERIAL_DTYPE

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "keepsake_amd" / "_lib" / "ks_render"


def bsdf_table(name, basecolor=(0.5, 0.5, 0.5), roughness=0.5, metallic=0.0, ior=1.5, r0=0.5, strans=0.0,
               dtrans=0.0, dfwd=0.0):
    def one(v):
        return f'{{ type = "constant", value = {v!r} }}'

    return f"""
[bsdf.{name}]
type = "principled_bsdf"
basecolor = {{ type = "constant", value = [{basecolor[0]!r}, {basecolor[1]!r}, {basecolor[2]!r}] }}
roughness = {one(roughness)}
metallic = {one(metallic)}
ior = {one(ior)}
specular_r0_mul = {one(r0)}
specular_trans = {one(strans)}
diffuse_trans = {one(dtrans)}
diffuse_trans_fwd = {one(dfwd)}
emissive = {{ type = "constant", value = [0.0, 0.0, 0.0] }}
"""


def task(object_name, materials, extra=""):
    mats = ", ".join(f'"{m}"' for m in materials)
    return f"""
[[task]]
type = "small_pt"
object = "{object_name}"
material = [[{mats}]]
bounces = 4
render_width = 32
render_height = 32
spp = 1
camera = {{ type = "perspective", pos = [0.0, 0.0, 5.0], target = [0.0, 0.0, 0.0], up = [0.0, 1.0, 0.0], vfov = 40.0, aspect = 1.0 }}
{extra}"""


MATS = bsdf_table("grey") + """
[material.grey]
bsdf = "bsdf.grey"
"""


def run(tmp_path, text, *args):
    cfg = tmp_path / "c.toml"
    cfg.write_text(f'output_dir = "{tmp_path / "out"}"\n' + text)
    return subprocess.run([str(TOOL), "-c", str(cfg), *args], capture_output=True, text=True)


def dump(tmp_path, text):
    d = tmp_path / "dump"
    r = run(tmp_path, text, "--dump-scene", str(d))
    assert r.returncode == 0, r.stderr
    info = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    t = d / "task_0"

    def rd(name, dt, cols=None):
        a = np.fromfile(t / name, dtype=dt)
        return a.reshape(-1, cols) if cols else a

    return info[0], {
        "vertices": rd("vertices.f32", "<f4", 3),
        "normals": rd("normals.f32", "<f4", 3),
        "indices": rd("indices.u32", "<u4", 3),
        "tri_geom": rd("tri_geom.u32", "<u4"),
        "geom_material": rd("geom_material.u32", "<u4"),
        "geom_flags": rd("geom_flags.u8", "u1"),
        "materials": rd("materials.bin", MATERIAL_DTYPE),
        "lights": rd("lights.bin", LIGHT_DTYPE),
    }


def asset(name, path, **kw):
    """A [mesh_asset.*] table; OBJ tables need load_materials (a required key, mesh_asset.cpp:249)."""
    if "format" not in kw:
        kw.setdefault("load_materials", "false")
    lines = [f"[mesh_asset.{name}]", f'path = "{path}"']
    for k, v in kw.items():
        lines.append(f"{k} = {v}")
    return "\n".join(lines) + "\n"


SHAPES_OBJ = """# three shapes
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 4 0 0
v 1 1 0
vn 0 0 1
vt 0.25 0.75
o square
f 1/1/1 2/1/1 3/1/1 4/1/1
o tri
f -3 1 2
g kite
f 1 6 7 4
"""


def test_obj_shapes_remap_and_quads(tmp_path):
    """Three shapes; quads split on the shorter diagonal; corners numbered in triangle order."""
    (tmp_path / "m.obj").write_text(SHAPES_OBJ)
    text = asset("m", tmp_path / "m.obj") + MATS + task("mesh_asset.m", ["material.grey"] * 3)
    info, s = dump(tmp_path, text)
    assert info["n_geoms"] == 3 and info["n_tris"] == 5 and info["n_materials"] == 1
    # square: |v2 - v0|^2 == |v3 - v1|^2 (2 == 2), not shorter, so [0,1,3],[1,2,3]: corners 1,2,4 then 3
    # tri: v5, v1, v2 (a negative index counts back from the vertices read so far)
    # kite (0,0,0) (4,0,0) (1,1,0) (0,1,0): |v2 - v0|^2 = 2 < |v3 - v1|^2 = 17, so [0,1,2],[0,2,3]
    want_v = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)]
    want_v += [(0, 0, 1), (0, 0, 0), (1, 0, 0)]
    want_v += [(0, 0, 0), (4, 0, 0), (1, 1, 0), (0, 1, 0)]
    assert np.array_equal(s["vertices"], np.array(want_v, np.float32))
    want_i = [(0, 1, 2), (1, 3, 2), (4, 5, 6), (7, 8, 9), (7, 9, 10)]
    assert np.array_equal(s["indices"], np.array(want_i, np.uint32))
    assert list(s["tri_geom"]) == [0, 0, 1, 2, 2]
    assert list(s["geom_material"]) == [0, 0, 0]
    # the square has vertex normals; the others take the geometric normal (flag) and zero rows
    assert np.array_equal(s["normals"][:4], np.tile([0, 0, 1], (4, 1)).astype(np.float32))
    assert not s["normals"][4:].any()
    # OBJ meshes default to one-sided (twosided = false); only the square has texture coordinates
    no_uv = 4  # KS_GEOM_NO_TEXCOORD
    assert list(s["geom_flags"]) == [KS_GEOM_ONE_SIDED, KS_GEOM_ONE_SIDED | KS_GEOM_FLAT | no_uv,
                                     KS_GEOM_ONE_SIDED | KS_GEOM_FLAT | no_uv]
    tc = np.fromfile(tmp_path / "dump" / "task_0" / "texcoords.f32", np.float32).reshape(-1, 2)
    assert np.array_equal(tc[:4], np.tile([0.25, 0.75], (4, 1)).astype(np.float32)) and not tc[4:].any()


def test_obj_twosided_flat_options(tmp_path):
    (tmp_path / "m.obj").write_text(SHAPES_OBJ)
    text = asset("m", tmp_path / "m.obj", twosided="true", use_smooth_normal="false") + MATS
    info, s = dump(tmp_path, text + task("mesh_asset.m", ["material.grey"] * 3))
    assert list(s["geom_flags"]) == [KS_GEOM_FLAT, KS_GEOM_FLAT | 4, KS_GEOM_FLAT | 4]
    assert s["normals"].size == 0  # no geometry keeps its vertex normals
    text = asset("m", tmp_path / "m.obj", twosided="true") + MATS
    info, s = dump(tmp_path, text + task("mesh_asset.m", ["material.grey"] * 3))
    assert list(s["geom_flags"]) == [0, KS_GEOM_FLAT | 4, KS_GEOM_FLAT | 4]  # 4: KS_GEOM_NO_TEXCOORD


def _poly_obj(pts):
    lines = [f"v {x!r} {y!r} {z!r}" for x, y, z in pts]
    lines.append("f " + " ".join(str(i + 1) for i in range(len(pts))))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize(
    "pts2",
    [
        [(0, 0), (2, 0), (3, 1.5), (1, 3), (-1, 1.5)],  # convex pentagon
        [(0, 0), (4, 0), (4, 1), (1, 1), (1, 3), (0, 3)],  # L-shaped hexagon (one reflex corner)
        [(0, 0), (4, 0), (4, 4), (2, 1), (0, 4)],  # deep notch
        [(0, 0), (3, 0), (3, 2), (2, 2), (2, 1), (1, 1), (1, 2), (0, 2)],  # U shape (two reflex corners)
    ],
)
@pytest.mark.parametrize("plane", ["xy", "zx", "tilted"])
def test_obj_polygon_triangulation(tmp_path, pts2, plane):
    """Polygons with more than four corners: the triangles tile the polygon exactly (areas sum to the
    shoelace area, every triangle keeps the polygon's orientation, n - 2 triangles)."""
    p2 = np.array(pts2, np.float64)
    area = 0.5 * np.sum(p2[:, 0] * np.roll(p2[:, 1], -1) - np.roll(p2[:, 0], -1) * p2[:, 1])
    if plane == "xy":
        pts = [(x, y, 0.0) for x, y in pts2]
    elif plane == "zx":
        pts = [(y, 1.0, x) for x, y in pts2]
    else:
        pts = [(x, 0.5 * y, 0.8660254037844386 * y) for x, y in pts2]
    (tmp_path / "p.obj").write_text(_poly_obj(pts))
    info, s = dump(tmp_path, asset("p", tmp_path / "p.obj") + MATS + task("mesh_asset.p", ["material.grey"]))
    n = len(pts2)
    assert info["n_tris"] == n - 2
    # map output vertices back to the 2D polygon corners
    v = s["vertices"].astype(np.float64)
    corner = [int(np.argmin(np.abs(np.array(pts) - x).sum(axis=1))) for x in v]
    tri2 = p2[np.array(corner)[s["indices"]]]
    a, b, c = tri2[:, 0], tri2[:, 1], tri2[:, 2]
    tri_area = 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    assert np.all(np.sign(tri_area) == np.sign(area))
    assert np.abs(tri_area).min() > 1e-9
    assert tri_area.sum() == pytest.approx(area, rel=1e-6)


@pytest.mark.parametrize(
    "body,msg",
    [
        ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2 3\n", "either all or no vertices have uv"),
        ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1 2 3//1\n", "either all or no vertices have a vertex normal"),
        ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "index 4 out of range"),
        ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "index 0 out of range"),
        ("v 0 0 0\nv 1 0 0\nf 1 2\n", "face with < 3 vertices"),
        ("v 0 0 0\n", "mesh asset has no triangles"),
    ],
)
def test_obj_errors(tmp_path, body, msg):
    (tmp_path / "e.obj").write_text(body)
    n_mats = 0 if "no triangles" in msg else 1
    r = run(tmp_path, asset("e", tmp_path / "e.obj") + MATS + task("mesh_asset.e", ["material.grey"] * n_mats),
            "--print-config")
    assert r.returncode != 0
    assert msg in r.stderr


def read_binary_asset(path):
    """The binary mesh asset layout, read field by field."""
    b = Path(path).read_bytes()
    magic = b"i_am_a_binary_mesh_asset"
    assert b[: len(magic)] == magic
    off = len(magic)
    (n,) = struct.unpack_from("<i", b, off)
    off += 4
    meshes = []
    for _ in range(n):
        ts, sn, nv, nt, tc, vn = struct.unpack_from("<??ii??", b, off)
        off += 12
        m = {"twosided": ts, "use_smooth_normal": sn}
        verts = np.frombuffer(b, "<f4", 3 * nv + 1, off)
        off += 4 * (3 * nv + 1)
        assert np.isnan(verts[-1])
        m["vertices"] = verts[:-1].reshape(-1, 3)
        m["indices"] = np.frombuffer(b, "<u4", 3 * nt, off).reshape(-1, 3)
        off += 12 * nt
        if tc:
            t = np.frombuffer(b, "<f4", 2 * nv + 2, off)
            off += 4 * (2 * nv + 2)
            m["texcoords"] = t[:-2].reshape(-1, 2)
        if vn:
            t = np.frombuffer(b, "<f4", 3 * nv + 1, off)
            off += 4 * (3 * nv + 1)
            m["vertex_normals"] = t[:-1].reshape(-1, 3)
        meshes.append(m)
    assert off == len(b)
    return meshes


def test_convert_and_load_binary(tmp_path):
    """convert_mesh_asset writes <task_dir>/<name>.bin; loading it gives the OBJ's scene back."""
    (tmp_path / "m.obj").write_text(SHAPES_OBJ)
    conv = asset("m", tmp_path / "m.obj", twosided="true") + """
[[task]]
type = "convert_mesh_asset"
task_dir = "conv"
assets = ["mesh_asset.m"]
"""
    r = run(tmp_path, conv)
    assert r.returncode == 0, r.stderr
    binf = tmp_path / "out" / "conv" / "m.bin"
    meshes = read_binary_asset(binf)
    assert len(meshes) == 3
    assert [m["twosided"] for m in meshes] == [True] * 3
    assert np.array_equal(meshes[0]["vertices"], np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], np.float32))
    assert np.array_equal(meshes[0]["indices"], np.array([(0, 1, 2), (1, 3, 2)], np.uint32))
    assert np.array_equal(meshes[0]["texcoords"], np.tile(np.float32([0.25, 0.75]), (4, 1)))
    assert "texcoords" not in meshes[1] and "vertex_normals" not in meshes[1]

    _, a = dump(tmp_path, asset("m", tmp_path / "m.obj", twosided="true") + MATS
                + task("mesh_asset.m", ["material.grey"] * 3))
    _, b = dump(tmp_path, asset("m", binf, format='"bin"') + MATS + task("mesh_asset.m", ["material.grey"] * 3))
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_binary_rejects_bad_files(tmp_path):
    (tmp_path / "bad.bin").write_bytes(b"not_a_binary_mesh_asset!" + b"\0" * 8)
    r = run(tmp_path, asset("b", tmp_path / "bad.bin", format='"bin"') + MATS + task("mesh_asset.b", ["material.grey"]),
            "--print-config")
    assert r.returncode != 0 and "invalid binary mesh asset" in r.stderr
    good = b"i_am_a_binary_mesh_asset" + struct.pack("<i??ii??", 1, True, True, 3, 1, False, False)
    good += np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, np.nan]).tobytes() + np.uint32([0, 1, 2]).tobytes()
    (tmp_path / "ok.bin").write_bytes(good)
    info, s = dump(tmp_path, asset("b", tmp_path / "ok.bin", format='"bin"') + MATS
                   + task("mesh_asset.b", ["material.grey"]))
    assert info["n_tris"] == 1 and info["n_vertices"] == 3
    (tmp_path / "trunc.bin").write_bytes(good[:-4])
    r = run(tmp_path, asset("b", tmp_path / "trunc.bin", format='"bin"') + MATS
            + task("mesh_asset.b", ["material.grey"]), "--print-config")
    assert r.returncode != 0 and "truncated" in r.stderr
    bad_idx = good[:-4] + np.uint32([3]).tobytes()
    (tmp_path / "idx.bin").write_bytes(bad_idx)
    r = run(tmp_path, asset("b", tmp_path / "idx.bin", format='"bin"') + MATS
            + task("mesh_asset.b", ["material.grey"]), "--print-config")
    assert r.returncode != 0 and "index out of range" in r.stderr


def test_to_world_transforms_points_and_normals(tmp_path):
    """Points by the affine matrix, normals by the inverse transpose of its 3x3 block, renormalised."""
    rng = np.random.default_rng(5)
    p = rng.normal(size=(6, 3)).astype(np.float32)
    n = rng.normal(size=(6, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    body = "".join(f"v {x!r} {y!r} {z!r}\n" for x, y, z in p.tolist())
    body += "".join(f"vn {x!r} {y!r} {z!r}\n" for x, y, z in n.tolist())
    body += "f 1//1 2//2 3//3\nf 4//4 5//5 6//6\n"
    (tmp_path / "t.obj").write_text(body)
    M = np.array([[2.0, 0.3, 0.0, 1.0], [0.0, 0.5, -0.2, 2.0], [0.1, 0.0, 3.0, -1.5], [0, 0, 0, 1]])
    flat = ", ".join(repr(float(x)) for x in M.reshape(-1))
    info, s = dump(tmp_path, asset("t", tmp_path / "t.obj", to_world=f"{{ matrix_rm = [{flat}] }}") + MATS
                   + task("mesh_asset.t", ["material.grey"]))
    M32 = M.astype(np.float32).astype(np.float64)
    want_p = p.astype(np.float64) @ M32[:3, :3].T + M32[:3, 3]
    want_n = n.astype(np.float64) @ np.linalg.inv(M32[:3, :3])
    want_n /= np.linalg.norm(want_n, axis=1, keepdims=True)
    np.testing.assert_allclose(s["vertices"], want_p, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(s["normals"], want_n, rtol=1e-6, atol=1e-6)
    # a rotation/translation/scale table is accepted too and leaves normals unit length
    info, s = dump(tmp_path, asset("t", tmp_path / "t.obj",
                                   to_world="{ scale = [1.0, 2.0, 0.5], rotation = { euler = [30.0, 10.0, -45.0] }, "
                                            "translation = [1.0, 0.0, 3.0] }")
                   + MATS + task("mesh_asset.t", ["material.grey"]))
    np.testing.assert_allclose(np.linalg.norm(s["normals"], axis=1), 1.0, atol=1e-6)


TRI_OBJ = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\no a\nf 1 2 3\no b\nf 1 3 4\no c\nf 1 2 4\n"


def test_material_tables(tmp_path):
    (tmp_path / "t.obj").write_text(TRI_OBJ)
    text = asset("t", tmp_path / "t.obj")
    text += bsdf_table("glass", basecolor=(1.0, 0.9, 0.8), roughness=0.05, ior=1.33, r0=0.25, strans=1.0)
    text += bsdf_table("cloth", basecolor=(0.2, 0.4, 0.6), roughness=0.9, metallic=0.25, dtrans=0.5, dfwd=0.75)
    text += """
[material.glass]
bsdf = "bsdf.glass"
[material.cloth]
bsdf = "bsdf.cloth"
[material.lamp]
bsdf = "bsdf.cloth"
emission = [10.0, 9.0, 8.0]
"""
    info, s = dump(tmp_path, text + task("mesh_asset.t", ["material.cloth", "material.glass", "material.cloth"]))
    assert list(s["geom_material"]) == [0, 1, 0]  # each material table is read once, in first-use order
    m = s["materials"]
    assert m.shape == (2,)
    assert np.array_equal(m[0]["basecolor"], np.float32([0.2, 0.4, 0.6]))
    assert m[0]["roughness"] == np.float32(0.9) and m[0]["metallic"] == np.float32(0.25)
    assert m[0]["diffuse_trans"] == np.float32(0.5) and m[0]["diffuse_trans_fwd"] == np.float32(0.75)
    assert m[1]["ior"] == np.float32(1.33) and m[1]["specular_r0_mul"] == np.float32(0.25)
    assert m[1]["specular_trans"] == np.float32(1.0)
    # the bsdf's own `emissive` field does not make a light; the material's `emission` does
    assert m[0]["has_emission"] == 0 and m[1]["has_emission"] == 0
    info, s = dump(tmp_path, text + task("mesh_asset.t", ["material.lamp", "material.glass", "material.cloth"]))
    m = s["materials"]
    assert m[0]["has_emission"] == 1 and np.array_equal(m[0]["emission"], np.float32([10, 9, 8]))
    assert m[2]["has_emission"] == 0


@pytest.mark.parametrize(
    "mats,extra_tables,msg",
    [
        (["material.grey"], "", "Material list mismatch dimension: 1 materials for 3 meshes"),
        (["material.grey", "material.grey", "material.nope"], "", "unknown asset 'material.nope'"),
        (["material.grey", "material.grey", "material.lamb"],
         '[bsdf.lamb]\ntype = "lambertian"\nalbedo = { type = "constant", value = [0.5, 0.5, 0.5] }\n'
         '[material.lamb]\nbsdf = "bsdf.lamb"\n', "bsdf type 'lambertian' is not supported"),
        (["material.grey", "material.grey", "material.tex"],
         bsdf_table("tex").replace('basecolor = { type = "constant", value = [0.5, 0.5, 0.5] }',
                                   'basecolor = { type = "texture", map = "texture.x" }')
         + '[material.tex]\nbsdf = "bsdf.tex"\n', "unknown asset 'texture.x'"),
        (["material.grey", "material.grey", "material.nm"],
         '[material.nm]\nbsdf = "bsdf.grey"\nnormal_map = "normal_map.x"\n', "unknown asset 'normal_map.x'"),
    ],
)
def test_material_errors(tmp_path, mats, extra_tables, msg):
    (tmp_path / "t.obj").write_text(TRI_OBJ)
    r = run(tmp_path, asset("t", tmp_path / "t.obj") + MATS + extra_tables + task("mesh_asset.t", mats),
            "--print-config")
    assert r.returncode != 0
    assert msg in r.stderr


def test_object_needs_material_list_and_camera(tmp_path):
    (tmp_path / "t.obj").write_text(TRI_OBJ)
    base = asset("t", tmp_path / "t.obj") + MATS + task("mesh_asset.t", ["material.grey"] * 3)
    r = run(tmp_path, base.replace('material = [["material.grey", "material.grey", "material.grey"]]\n', ""),
            "--print-config")
    assert r.returncode != 0 and "No materials" in r.stderr
    cam = base[base.index("camera = "):].splitlines()[0]
    r = run(tmp_path, base.replace(cam, ""), "--print-config")
    assert r.returncode != 0 and "camera" in r.stderr
    r = run(tmp_path, asset("t", tmp_path / "t.obj", load_materials="true") + MATS
            + task("mesh_asset.t", ["material.grey"] * 3), "--print-config")
    assert r.returncode != 0 and "load_materials" in r.stderr
    r = run(tmp_path, asset("t", tmp_path / "t.obj", format='"fbx"') + MATS
            + task("mesh_asset.t", ["material.grey"] * 3), "--print-config")
    assert r.returncode != 0 and "Unsupported mesh asset format [fbx]" in r.stderr
