"""glTF compound mesh assets on the TOML path (keepsake_amd/host/gltf_asset.cpp), checked without a GPU
through `ks_render --dump-scene`: buffers as data URIs, external files and .glb chunks; strided and
16/32-bit index accessors; the node hierarchy (world = parent * local, node.matrix column-major or
T * R * S, traverse_gltf_scene_graph mesh_asset.cpp:486-534) flattened per instance; the material
conversion of CompoundMeshAsset::load_from_gltf (mesh_asset.cpp:650-956). Expected world coordinates
come from numpy matrix products, independent of the loader's own arithmetic."""

import base64
import json
import struct

import numpy as np
import pytest

from keepsake_amd.api import KS_GEOM_FLAT, KS_GEOM_ONE_SIDED
from test_mesh_asset import MATS, bsdf_table, dump, run


class Gltf:
    """A tiny glTF writer: one buffer, a view per accessor."""

    def __init__(self):
        self.blob = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                    "nodes": [], "scenes": [{"nodes": []}], "materials": []}

    def accessor(self, arr, comp, typ, stride=None):
        arr = np.ascontiguousarray(arr)
        while len(self.blob) % 4:
            self.blob.append(0)
        off = len(self.blob)
        raw = arr.tobytes()
        n = arr.shape[0]
        elem = raw and len(raw) // n
        view = {"buffer": 0, "byteOffset": off}
        if stride and stride > elem:  # interleave with padding bytes
            out = bytearray()
            for i in range(n):
                out += raw[i * elem:(i + 1) * elem] + b"\xab" * (stride - elem)
            raw = bytes(out)
            view["byteStride"] = stride
        self.blob += raw
        view["byteLength"] = len(raw)
        self.doc["bufferViews"].append(view)
        self.doc["accessors"].append({"bufferView": len(self.doc["bufferViews"]) - 1, "componentType": comp,
                                      "count": n, "type": typ})
        return len(self.doc["accessors"]) - 1

    def primitive(self, pos, idx=None, normals=None, index_type=5125, stride=None, material=None):
        p = {"attributes": {"POSITION": self.accessor(np.float32(pos), 5126, "VEC3", stride)}}
        if normals is not None:
            p["attributes"]["NORMAL"] = self.accessor(np.float32(normals), 5126, "VEC3")
        if idx is not None:
            dt = {5125: np.uint32, 5123: np.uint16, 5121: np.uint8}[index_type]
            p["indices"] = self.accessor(np.asarray(idx, dt).reshape(-1), index_type, "SCALAR")
        if material is not None:
            p["material"] = material
        return p

    def mesh(self, prims):
        self.doc["meshes"].append({"primitives": prims})
        return len(self.doc["meshes"]) - 1

    def node(self, root=False, **kw):
        self.doc["nodes"].append(kw)
        i = len(self.doc["nodes"]) - 1
        if root:
            self.doc["scenes"][0]["nodes"].append(i)
        return i

    def material(self, **kw):
        self.doc["materials"].append(kw)
        return len(self.doc["materials"]) - 1

    def write_gltf(self, path, external=False):
        doc = json.loads(json.dumps(self.doc))
        if external:
            (path.parent / "buf.bin").write_bytes(bytes(self.blob))
            doc["buffers"] = [{"uri": "buf.bin", "byteLength": len(self.blob)}]
        else:
            uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.blob)).decode()
            doc["buffers"] = [{"uri": uri, "byteLength": len(self.blob)}]
        path.write_text(json.dumps(doc, indent=1))

    def write_glb(self, path):
        doc = json.loads(json.dumps(self.doc))
        doc["buffers"] = [{"byteLength": len(self.blob)}]
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        bin_ = bytes(self.blob) + b"\0" * (-len(self.blob) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bin_), 0x004E4942) + bin_
        path.write_bytes(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)


def compound(name, path, **kw):
    kw.setdefault("load_materials", "false")
    lines = [f"[compound_mesh_asset.{name}]", f'path = "{path}"'] + [f"{k} = {v}" for k, v in kw.items()]
    return "\n".join(lines) + "\n"


def task(name, materials=None):
    mat = ""
    if materials is not None:
        mat = "material = [" + ", ".join("[" + ", ".join(f'"{m}"' for m in row) + "]" for row in materials) + "]\n"
    return f"""
[[task]]
type = "small_pt"
compound_object = "{name}"
{mat}bounces = 4
render_width = 32
render_height = 32
spp = 1
camera = {{ type = "perspective", pos = [0.0, 0.0, 5.0], target = [0.0, 0.0, 0.0], up = [0.0, 1.0, 0.0], vfov = 40.0, aspect = 1.0 }}
"""


def trs(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(s)
    M[:3, 3] = t
    return M


TRI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
QUAD = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], np.float32)
QUAD_IDX = [0, 1, 2, 0, 2, 3]


def scene_graph():
    g = Gltf()
    rng = np.random.default_rng(3)
    nrm = rng.normal(size=(4, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    m0 = g.mesh([g.primitive(QUAD, QUAD_IDX, normals=nrm, index_type=5123, stride=20),
                 g.primitive(TRI, [0, 1, 2])])
    m1 = g.mesh([g.primitive(TRI * 3, None)])  # non-indexed
    q = np.array([0.2, -0.4, 0.1, 0.9])
    q = tuple(q / np.linalg.norm(q))
    child_m = trs(t=(0.5, -1, 2), q=(0, 0.6, 0, 0.8), s=(1, 2, 1))
    child = g.node(matrix=list(child_m.T.reshape(-1)), mesh=m0)  # column-major
    g.node(root=True, translation=[1.0, 2.0, 3.0], rotation=list(q), scale=[2.0, 2.0, 0.5], children=[child])
    g.node(root=True, mesh=m1, translation=[0.0, 0.0, -4.0])
    g.node(root=True, mesh=m0)
    world = [trs((1, 2, 3), q, (2, 2, 0.5)) @ child_m, trs((0, 0, -4)), np.eye(4)]
    return g, nrm, world


def _expect(world, nrm):
    """Per geometry (instance order, then primitive order): world vertices and normals."""
    geoms = []
    for inst, (mesh, M) in enumerate(zip([0, 1, 0], world)):
        prims = [(QUAD, nrm), (TRI, None)] if mesh == 0 else [(TRI * 3, None)]
        for pos, n in prims:
            p = pos.astype(np.float64) @ M[:3, :3].T + M[:3, 3]
            wn = None
            if n is not None:
                wn = n @ np.linalg.inv(M[:3, :3])
                wn /= np.linalg.norm(wn, axis=1, keepdims=True)
            geoms.append((p, wn))
    return geoms


@pytest.mark.parametrize("container", ["data_uri", "external", "glb"])
def test_scene_graph_flattened(tmp_path, container):
    g, nrm, world = scene_graph()
    path = tmp_path / ("s.glb" if container == "glb" else "s.gltf")
    if container == "glb":
        g.write_glb(path)
    else:
        g.write_gltf(path, external=container == "external")
    mats = [["material.grey", "material.grey"], ["material.grey"]]
    info, s = dump(tmp_path, compound("s", path) + MATS + task("compound_mesh_asset.s", mats))
    assert info["n_geoms"] == 5 and info["n_tris"] == 2 + 1 + 1 + 2 + 1
    exp = _expect(world, nrm)
    vo = 0
    to = 0
    tri_counts = [2, 1, 1, 2, 1]
    for gi, ((p, wn), nt) in enumerate(zip(exp, tri_counts)):
        nv = p.shape[0]
        np.testing.assert_allclose(s["vertices"][vo:vo + nv], p, rtol=1e-5, atol=2e-5)
        if wn is not None:
            np.testing.assert_allclose(s["normals"][vo:vo + nv], wn, rtol=1e-5, atol=2e-5)
        else:
            assert not s["normals"][vo:vo + nv].any()
        local = s["indices"][to:to + nt] - vo
        want = np.array(QUAD_IDX if nt == 2 else [0, 1, 2]).reshape(-1, 3)
        assert np.array_equal(local, want), gi
        assert np.all(s["tri_geom"][to:to + nt] == gi)
        vo += nv
        to += nt
    # glTF meshes are two-sided; only the quads carry vertex normals
    assert list(s["geom_flags"]) == [0, KS_GEOM_FLAT, KS_GEOM_FLAT, 0, KS_GEOM_FLAT]
    assert list(s["geom_material"]) == [0] * 5


def test_material_list_dimensions(tmp_path):
    g, _, _ = scene_graph()
    g.write_gltf(tmp_path / "s.gltf")
    r = run(tmp_path, compound("s", tmp_path / "s.gltf") + MATS + task("compound_mesh_asset.s", [["material.grey"]]),
            "--print-config")
    assert r.returncode != 0 and "1 lists for 2 subscenes" in r.stderr
    r = run(tmp_path, compound("s", tmp_path / "s.gltf") + MATS
            + task("compound_mesh_asset.s", [["material.grey"], ["material.grey"]]), "--print-config")
    assert r.returncode != 0 and "1 materials for 2 primitives of mesh 0" in r.stderr
    r = run(tmp_path, compound("s", tmp_path / "s.gltf") + MATS + task("compound_mesh_asset.s"), "--print-config")
    assert r.returncode != 0 and "No materials" in r.stderr


def _material_scene(**mat):
    g = Gltf()
    m = g.material(**mat) if mat else None
    g.mesh([g.primitive(TRI, [0, 1, 2], material=m)])
    g.node(root=True, mesh=0)
    return g


@pytest.mark.parametrize(
    "mat,want,flags",
    [
        ({}, dict(basecolor=[1, 1, 1], roughness=1.0, metallic=1.0, ior=1.5, specular_r0_mul=0.5, specular_trans=0.0,
                  has_emission=0), 0),  # glTF's default material
        ({"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 1.0], "metallicFactor": 0.25,
                                   "roughnessFactor": 0.75}},
         dict(basecolor=[0.2, 0.4, 0.6], roughness=0.75, metallic=0.25), 0),
        ({"emissiveFactor": [1.0, 0.5, 0.25], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 8}}},
         dict(emission=[8.0, 4.0, 2.0], has_emission=1), 0),
        ({"emissiveFactor": [1.0, 0.5, 0.25], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 0}}},
         dict(emission=[0.0, 0.0, 0.0], has_emission=0), 0),
        ({"extensions": {"KHR_materials_ior": {"ior": 1.33}}}, dict(ior=1.33), 0),
        ({"extensions": {"KHR_materials_specular": {"specularFactor": 0.3}}}, dict(specular_r0_mul=0.3), 0),
        ({"extensions": {"KHR_materials_specular": {"specularColorFactor": [0.5, 1.0, 0.0]}}},
         dict(specular_r0_mul=0.5), 0),
        ({"extensions": {"KHR_materials_specular": {"specularFactor": 0.5, "specularColorFactor": [1.0, 0.5, 0.0]}}},
         dict(specular_r0_mul=0.25), 0),
        ({"extensions": {"KHR_materials_specular": {}}}, dict(specular_r0_mul=0.5), 0),
        ({"extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.9}}}, dict(specular_trans=0.9),
         KS_GEOM_ONE_SIDED),
        ({"alphaMode": "MASK", "pbrMetallicRoughness": {"baseColorFactor": [1.0, 1.0, 1.0, 1.0]}}, dict(), 0),
    ],
)
def test_material_conversion(tmp_path, mat, want, flags):
    _material_scene(**mat).write_gltf(tmp_path / "m.gltf")
    info, s = dump(tmp_path, compound("m", tmp_path / "m.gltf", load_materials="true") + task("compound_mesh_asset.m"))
    m = s["materials"]
    assert m.shape == (1,)
    for k, v in want.items():
        np.testing.assert_array_equal(m[0][k], np.asarray(v, np.float32 if k != "has_emission" else np.int32), k)
    assert m[0]["diffuse_trans"] == 0 and m[0]["diffuse_trans_fwd"] == 0
    assert list(s["geom_flags"]) == [flags | KS_GEOM_FLAT]


def test_toml_material_list_overrides_gltf(tmp_path):
    _material_scene(extensions={"KHR_materials_transmission": {"transmissionFactor": 1.0}}).write_gltf(
        tmp_path / "m.gltf")
    text = compound("m", tmp_path / "m.gltf", load_materials="true") + MATS
    info, s = dump(tmp_path, text + task("compound_mesh_asset.m", [["material.grey"]]))
    assert s["materials"][0]["specular_trans"] == 0 and s["materials"][0]["roughness"] == np.float32(0.5)
    assert list(s["geom_flags"]) == [KS_GEOM_ONE_SIDED | KS_GEOM_FLAT]  # the transmission mesh stays one-sided


def _textured_gltf(tmp_path):
    """A triangle with TEXCOORD_0 and a material using every texture slot: images as a file, a data URI
    and a buffer view (8-bit RGB, 16-bit gray, 8-bit RGBA), samplers with and without mipmap filtering."""
    import image_files as imf

    g = Gltf()
    rgb = np.random.default_rng(0).integers(0, 256, (4, 5, 3)).astype(np.uint8)
    gray16 = np.random.default_rng(1).integers(0, 65536, (3, 3, 1)).astype(np.uint16)
    rgba = np.random.default_rng(2).integers(0, 256, (2, 2, 4)).astype(np.uint8)
    imf.write_png_any(tmp_path / "rgb.png", rgb, 2, 8)
    imf.write_png_any(tmp_path / "g16.png", gray16, 0, 16)
    imf.write_png_any(tmp_path / "rgba.png", rgba, 6, 8)
    while len(g.blob) % 4:
        g.blob.append(0)
    off = len(g.blob)
    png = (tmp_path / "rgba.png").read_bytes()
    g.blob += png
    g.doc["bufferViews"].append({"buffer": 0, "byteOffset": off, "byteLength": len(png)})
    data_uri = "data:image/png;base64," + base64.b64encode((tmp_path / "g16.png").read_bytes()).decode()
    g.doc["images"] = [{"uri": "rgb.png"}, {"uri": data_uri}, {"bufferView": len(g.doc["bufferViews"]) - 1,
                                                              "mimeType": "image/png"}]
    g.doc["samplers"] = [{"minFilter": 9987}, {"minFilter": 9729}]
    g.doc["textures"] = [{"source": 0, "sampler": 0}, {"source": 1, "sampler": 1}, {"source": 2}]
    tt = {"extensions": {"KHR_texture_transform": {"scale": [2.0, 3.0], "offset": [0.25, 0.5]}}}
    m = g.material(
        pbrMetallicRoughness={"baseColorTexture": {"index": 2, **tt}, "baseColorFactor": [0.5, 0.6, 0.7, 0.8],
                              "metallicRoughnessTexture": {"index": 1}, "metallicFactor": 0.9, "roughnessFactor": 0.4},
        normalTexture={"index": 0, "scale": 0.7}, emissiveTexture={"index": 0}, emissiveFactor=[1.0, 0.5, 0.25],
        alphaMode="BLEND",
        extensions={"KHR_materials_specular": {"specularFactor": 0.8, "specularTexture": {"index": 1}},
                    "KHR_materials_transmission": {"transmissionFactor": 0.3, "transmissionTexture": {"index": 2}}})
    p = g.primitive(TRI, [0, 1, 2], material=m)
    p["attributes"]["TEXCOORD_0"] = g.accessor(np.float32([[0, 0], [1, 0], [0, 1]]), 5126, "VEC2")
    g.mesh([p])
    g.node(root=True, mesh=0)
    g.write_gltf(tmp_path / "m.gltf")
    return rgb, gray16, rgba


def test_gltf_textures(tmp_path):
    import ctypes as C

    from keepsake_amd import api

    rgb, gray16, rgba = _textured_gltf(tmp_path)
    info, s = dump(tmp_path, compound("m", tmp_path / "m.gltf", load_materials="true") + task("compound_mesh_asset.m"))
    d = tmp_path / "dump" / "task_0"
    assert np.fromfile(d / "field_of.i32", np.int32).tolist() == [0]
    f = api.MaterialFields.from_buffer_copy((d / "fields.bin").read_bytes())
    np.testing.assert_array_equal(np.fromfile(d / "texcoords.f32", np.float32).reshape(-1, 2), [[0, 0], [1, 0], [0, 1]])
    lines = [list(map(int, l.split())) for l in (d / "textures.txt").read_text().split("\n") if l.strip()]
    # images in first-use order: rgba (base colour), gray16 (metallic-roughness), rgb (normal / emissive)
    assert lines == [[2, 2, 4, api.KS_TEX_U8, api.KS_COLORSPACE_SRGB, 1, 1],
                     [3, 3, 4, api.KS_TEX_U16, api.KS_COLORSPACE_LINEAR, 1, 1],
                     [5, 4, 4, api.KS_TEX_U8, api.KS_COLORSPACE_SRGB, 1, 1]]
    np.testing.assert_array_equal(np.fromfile(d / "texture_0.bin", np.uint8).reshape(2, 2, 4), rgba)
    g16 = np.fromfile(d / "texture_1.bin", np.uint16).reshape(3, 3, 4)
    np.testing.assert_array_equal(g16[..., :3], np.repeat(gray16, 3, -1))
    assert (g16[..., 3] == 65535).all()
    t2 = np.fromfile(d / "texture_2.bin", np.uint8).reshape(4, 5, 4)
    np.testing.assert_array_equal(t2[..., :3], rgb)
    assert (t2[..., 3] == 255).all()
    bc = f.basecolor
    assert bc.texture == 0 and bc.sampler == api.KS_SAMPLER_NEAREST and bc.flip_v == 0  # texture 2 has no sampler
    np.testing.assert_allclose(list(bc.scale)[:3], [0.5, 0.6, 0.7])
    assert list(bc.uv_scale) == [2.0, 3.0] and list(bc.uv_offset) == [0.25, 0.5]
    assert f.has_opacity_map == 1 and f.opacity.texture == 0 and f.opacity.swizzle[0] == 3
    assert f.opacity.scale[0] == np.float32(0.8)
    assert (f.roughness.texture, f.roughness.swizzle[0], f.metallic.swizzle[0]) == (1, 1, 2)
    assert f.roughness.sampler == api.KS_SAMPLER_NEAREST  # minFilter LINEAR: no mipmaps
    assert f.roughness.scale[0] == np.float32(0.4) and f.metallic.scale[0] == np.float32(0.9)
    assert f.has_normal_map == 1 and f.normal_map.texture == 2 and f.normal_map.sampler == api.KS_SAMPLER_EWA
    assert f.normal_strength == np.float32(0.7) and f.normal_range01 == 1
    assert f.has_emission == 1 and f.emission.texture == 2 and list(f.emission.scale)[:3] == [1.0, 0.5, 0.25]
    assert f.specular_r0_mul.texture == 1 and f.specular_r0_mul.scale[0] == np.float32(0.8)
    assert f.specular_trans.texture == 0 and f.specular_trans.scale[0] == np.float32(0.3)
    assert list(np.fromfile(d / "geom_flags.u8", np.uint8)) == [KS_GEOM_ONE_SIDED | KS_GEOM_FLAT]
    # ignore_textures keeps the constant factors only
    info, s = dump(tmp_path, compound("m", tmp_path / "m.gltf", load_materials="true", ignore_textures="true")
                   + task("compound_mesh_asset.m"))
    assert np.fromfile(tmp_path / "dump" / "task_0" / "field_of.i32", np.int32).tolist() == [0]
    f2 = api.MaterialFields.from_buffer_copy((tmp_path / "dump" / "task_0" / "fields.bin").read_bytes())
    assert f2.basecolor.texture == -1 and f2.has_opacity_map == 1 and f2.opacity.texture == -1  # alpha stays
    assert (tmp_path / "dump" / "task_0" / "textures.txt").read_text() == ""


def test_gltf_jpeg_image_is_an_error(tmp_path):
    g = _material_scene(pbrMetallicRoughness={"baseColorTexture": {"index": 0}})
    g.doc["images"] = [{"uri": "data:image/jpeg;base64," + base64.b64encode(b"\xff\xd8\xff\xe0 not decoded").decode()}]
    g.doc["textures"] = [{"source": 0}]
    g.write_gltf(tmp_path / "m.gltf")
    r = run(tmp_path, compound("m", tmp_path / "m.gltf", load_materials="true") + task("compound_mesh_asset.m"),
            "--print-config")
    assert r.returncode != 0 and "only PNG images" in r.stderr


def _broken(kind, tmp_path):
    g = Gltf()
    if kind == "mode":
        p = g.primitive(TRI, [0, 1, 2])
        p["mode"] = 1
    elif kind == "ubyte":
        p = g.primitive(TRI, [0, 1, 2], index_type=5121)
    elif kind == "index_range":
        p = g.primitive(TRI, [0, 1, 3])
    else:
        p = g.primitive(TRI, [0, 1, 2])
    g.mesh([p])
    if kind == "cycle":
        g.node(root=True, mesh=0, children=[1])
        g.node(children=[0])
    else:
        g.node(root=True, mesh=0)
    if kind == "overrun":
        g.doc["accessors"][0]["count"] = 4
    if kind == "version":
        g.doc["asset"]["version"] = "1.0"
    path = tmp_path / "b.gltf"
    g.write_gltf(path)
    if kind == "json":
        path.write_text(path.read_text()[:-3])
    if kind == "short_buffer":
        doc = json.loads(path.read_text())
        doc["buffers"][0]["byteLength"] += 8
        path.write_text(json.dumps(doc))
    return path


@pytest.mark.parametrize(
    "kind,msg",
    [
        ("mode", "Only support triangle primitives"),
        ("ubyte", "Unsupported index buffer component type"),
        ("index_range", "vertex index out of range"),
        ("cycle", "cycle"),
        ("overrun", "accessor runs past its bufferView"),
        ("version", "asset.version must be 2.x"),
        ("json", "JSON error"),
        ("short_buffer", "shorter than its byteLength"),
    ],
)
def test_malformed_files(tmp_path, kind, msg):
    path = _broken(kind, tmp_path)
    r = run(tmp_path, compound("b", path) + MATS + task("compound_mesh_asset.b", [["material.grey"]]),
            "--print-config")
    assert r.returncode != 0
    assert msg in r.stderr


def test_json_strings_and_numbers(tmp_path):
    """Escapes, unicode, exponents and whitespace in the JSON reader."""
    g = _material_scene(name="café \"quoted\" \\ 😀", pbrMetallicRoughness={"roughnessFactor": 2.5e-1})
    g.doc["extras"] = {"a": [True, False, None, -1.5E+2, {"b": "\t\n"}]}
    g.write_gltf(tmp_path / "m.gltf")
    text = (tmp_path / "m.gltf").read_text()
    (tmp_path / "m.gltf").write_text(text.replace("roughnessFactor\": 0.25", "roughnessFactor\" :\r\n 25e-2"))
    info, s = dump(tmp_path, compound("m", tmp_path / "m.gltf", load_materials="true") + task("compound_mesh_asset.m"))
    assert s["materials"][0]["roughness"] == np.float32(0.25)


def cornell_gltf(hs):
    """A flat scene (HostScene) as one glTF mesh: a primitive per geometry with the vertices it uses, a
    glTF material per geometry carrying its constant factors, the emission as emissiveFactor."""
    g = Gltf()
    prims = []
    for geom, mi in enumerate(hs.geom_material):
        m = hs.materials[int(mi)]
        tris = hs.indices[hs.tri_geom == geom]
        used, local = np.unique(tris.reshape(-1), return_inverse=True)
        mat = {"pbrMetallicRoughness": {"baseColorFactor": [float(x) for x in m["basecolor"]] + [1.0],
                                        "metallicFactor": float(m["metallic"]),
                                        "roughnessFactor": float(m["roughness"])},
               "extensions": {"KHR_materials_ior": {"ior": float(m["ior"])},
                              "KHR_materials_specular": {"specularFactor": float(m["specular_r0_mul"])}}}
        assert m["specular_trans"] == 0 and m["diffuse_trans"] == 0
        if m["has_emission"]:
            mat["emissiveFactor"] = [float(x) for x in m["emission"]]
        prims.append(g.primitive(hs.vertices[used], local.reshape(-1, 3), material=g.material(**mat)))
    g.mesh(prims)
    g.node(root=True, mesh=0)
    return g


def test_cornell_gltf_resolves_to_the_procedural_scene(tmp_path):
    """Triangle for triangle the same corner positions and material values as the procedural box."""
    from keepsake_amd import api

    hs = api.procedural_scene("cornell")
    cornell_gltf(hs).write_glb(tmp_path / "c.glb")
    info, s = dump(tmp_path, compound("c", tmp_path / "c.glb", load_materials="true") + task("compound_mesh_asset.c"))
    assert info["n_tris"] == hs.n_tris
    assert np.array_equal(s["vertices"][s["indices"]], hs.vertices[hs.indices])
    got = s["materials"][s["geom_material"][s["tri_geom"]]]
    want = hs.materials[hs.geom_material[hs.tri_geom]]
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
