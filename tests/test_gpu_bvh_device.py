"""The device BVH builder (KS_BVH_BUILD_DEVICE: clustering in Morton order): the trees it lays out, and the
hits they give.

Structure (read back with ks_scene_bvh_get): every triangle is referenced exactly once with its own
vertices, every child box contains what lies below it, leaves hold at most max_prims_in_node
triangles, and the leaf depth equals the reported max_depth (which sizes the traversal stack).

Hits: closest-hit and occlusion queries agree with the oracle's traversal, whose closest hits are pinned
against brute force in test_oracle_*; a render over the device tree equals the render over the host
SBVH except where two triangles tie at the same distance (the two trees test them in different
orders), and alpha-tested triangles keep their filter.
"""

import numpy as np
import pytest

from texture_scenes import textured_box

pytestmark = pytest.mark.gpu

LEAF = 0x80000000


def _log(*a):
    print(*a, flush=True)


def _children(node, width):
    """(lo (k, 3), hi (k, 3), child pointers) of one node's used slots."""
    f = node.view(np.float32)
    if width == 4:
        lo = np.stack([f[0:4], f[8:12], f[16:20]], 1)
        hi = np.stack([f[4:8], f[12:16], f[20:24]], 1)
        ch = node[24:28]
    else:
        lo = np.array([[f[0], f[2], f[8]], [f[4], f[6], f[10]]], np.float32)
        hi = np.array([[f[1], f[3], f[9]], [f[5], f[7], f[11]]], np.float32)
        ch = node[12:14]
    used = np.all(lo <= hi, axis=1)
    return lo[used], hi[used], ch[used]


BUILDERS = [1]  # KS_BVH_BUILD_DEVICE


def check_structure(sc, hs, builder, max_leaf=4):
    info = sc.info()
    nodes, refs = sc.bvh()
    width = info["branching"]
    n = hs.n_tris
    assert info["builder"] == builder and info["n_refs"] == n and info["n_tris"] == n
    tri = refs[:, 3].astype(np.int64)
    assert np.array_equal(np.sort(tri), np.arange(n)), "each triangle once"
    rf = refs.view(np.float32)
    v = np.asarray(hs.vertices, np.float32)[np.asarray(hs.indices, np.int64)[tri]]  # (n, 3, 3)
    assert np.array_equal(rf[:, 0:3], v[:, 0]) and np.array_equal(rf[:, 4:7], v[:, 1])
    assert np.array_equal(rf[:, 8:11], v[:, 2])
    covered = np.zeros(n, np.int32)
    max_depth = 0
    seen = np.zeros(len(nodes), bool)
    stack = [(0, 0)]
    while stack:
        ni, depth = stack.pop()
        assert not seen[ni], "node reached twice"
        seen[ni] = True
        lo, hi, ch = _children(nodes[ni], width)
        assert len(ch) >= 1
        for k in range(len(ch)):
            c = int(ch[k])
            if c & LEAF:
                first, count = (c >> 4) & 0x7FFFFFF, (c & 15) + 1
                assert count <= max_leaf and first + count <= n
                covered[first:first + count] += 1
                tv = v[first:first + count].reshape(-1, 3)
                assert np.all(tv >= lo[k]) and np.all(tv <= hi[k]), "leaf box holds its triangles"
                max_depth = max(max_depth, depth + 1)
            else:
                assert 0 < c < len(nodes)
                clo, chi, _ = _children(nodes[c], width)
                assert np.all(clo >= lo[k]) and np.all(chi <= hi[k]), "child box inside its slot"
                stack.append((c, depth + 1))
    assert seen.all(), "every node reachable"
    assert np.all(covered == 1), "every reference in exactly one leaf"
    assert max_depth == info["max_depth"]
    all_v = v.reshape(-1, 3)
    assert np.array_equal(np.float32(info["bound_lo"]), all_v.min(0))
    assert np.array_equal(np.float32(info["bound_hi"]), all_v.max(0))
    return info


def _compare_hits(g, o):
    gp, op = g[:, 3].view(np.int32), o[:, 3].view(np.int32)
    same = gp == op
    both = (gp >= 0) & (op >= 0)
    tie = both & (np.abs(g[:, 0] - o[:, 0]) <= 1e-6 * np.maximum(1.0, np.abs(o[:, 0])))
    agree = (same | tie).mean()
    hit = both & same
    trel = np.abs(g[hit, 0] - o[hit, 0]) / np.maximum(np.abs(o[hit, 0]), 1e-30)
    return agree, (trel.max() if trel.size else 0.0)


def _random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, d
    rays[:, 7] = rng.uniform(0.5, 2.5 * (hi - lo), n).astype(np.float32)
    return rays


@pytest.mark.parametrize("builder", BUILDERS)
@pytest.mark.parametrize("branching", [4, 2])
@pytest.mark.parametrize("which", ["cornell", "spheres_small"])
def test_device_bvh_structure_and_hits(device, ks, ora, cornell, spheres_small, which, branching, builder):
    hs = cornell if which == "cornell" else spheres_small
    _log("device bvh", which, "branching", branching, "builder", builder, "tris", hs.n_tris)
    sc = ks.Scene(device, hs, bvh={"builder": builder, "branching": branching})
    info = check_structure(sc, hs, builder)
    _log("nodes", info["n_nodes"], "depth", info["max_depth"], "build ms", info["bvh_build_ms"])
    host = ks.Scene(device, hs, bvh={"branching": branching})
    assert info["bound_lo"] == host.info()["bound_lo"] and info["bound_hi"] == host.info()["bound_hi"]
    osc = ora.OracleScene.from_host_scene(hs)
    lo, hi = (-0.2, 1.2) if which == "cornell" else (-12.0, 12.0)
    rays = _random_rays(100000, lo, hi, 7)
    g = sc.intersect(rays)
    o = osc.intersect(rays)
    agree, trel = _compare_hits(g, o)
    _log("random-ray agreement", agree, "t rel", trel)
    assert agree >= 0.9999 and trel <= 1e-5
    assert np.array_equal(sc.occluded(rays).astype(bool), o[:, 3].view(np.int32) >= 0)
    # against the host SBVH on the device: the same watertight test, so the same triangle gives the same
    # bits; the two trees may pick different triangles only where two tie at one distance (shared edges)
    h = host.intersect(rays)
    agree_h, _ = _compare_hits(g, h)
    same = h[:, 3].view(np.int32) == g[:, 3].view(np.int32)
    _log("agreement with host SBVH", agree_h, "same triangle", same.mean())
    assert agree_h >= 0.9999 and np.array_equal(h[same, :3], g[same, :3])
    host.close()
    sc.close()


def _soup(ks, tris):
    """A one-material scene of the given (T, 3, 3) triangles."""
    tris = np.asarray(tris, np.float32)
    t = len(tris)
    cam = dict(pos=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), vfov=40.0, aspect=1.0)
    mats = np.zeros(1, ks.MATERIAL_DTYPE)
    mats["basecolor"] = 0.5
    mats["ior"] = 1.5
    return ks.HostScene(tris.reshape(-1, 3), np.arange(3 * t, dtype=np.uint32).reshape(t, 3),
                        np.zeros(t, np.uint32), np.zeros(1, np.uint32), mats, np.zeros(0, ks.LIGHT_DTYPE), cam)


@pytest.mark.parametrize("case", ["one", "two", "stacked", "degenerate", "dup_centroids", "identical"])
def test_device_bvh_small_scenes(device, ks, ora, case):
    rng = np.random.default_rng(11)
    base = np.float32([[-1, -1, 0], [1, -1, 0], [0, 1, 0]])
    if case == "one":
        tris = [base]
    elif case == "two":
        tris = [base, base + np.float32([0.5, 0.0, -1.0])]
    elif case == "stacked":  # 40 parallel triangles one behind the other
        tris = [base + np.float32([0, 0, -0.1 * k]) for k in range(40)]
    elif case == "degenerate":  # a line and a point among ordinary triangles
        tris = [base, np.float32([[0, 0, 1], [1, 1, 1], [2, 2, 1]]), np.float32([[0.2, 0.2, 0.5]] * 3),
                base * 0.5 + np.float32([0, 0, 0.7])]
    elif case == "identical":  # 6000 copies of one triangle: every box is equal (clustering ties)
        tris = [base] * 6000
    else:  # 37 triangles whose boxes share one centre: every Morton key is equal
        tris = []
        for _ in range(37):
            a = rng.normal(size=3).astype(np.float32)
            b = (a * rng.uniform(-1, 1, 3)).astype(np.float32)  # inside the box spanned by a and -a
            tris.append(np.stack([a, -a, b]))
    hs = _soup(ks, tris)
    _log("small scene", case, hs.n_tris)
    for branching, builder in [(w, b) for w in (4, 2) for b in BUILDERS]:
        sc = ks.Scene(device, hs, bvh={"builder": builder, "branching": branching})
        check_structure(sc, hs, builder)
        rays = _random_rays(20000, -3.0, 3.0, 5)
        g = sc.intersect(rays)
        o = ora.OracleScene.from_host_scene(hs).intersect(rays)
        agree, trel = _compare_hits(g, o)
        assert agree >= 0.9999 and trel <= 1e-5, (case, branching, builder, agree, trel)
        sc.close()


def test_device_bvh_render_matches_host_sbvh(device, ks, cornell):
    W = H = 128
    p = ks.small_pt_params(bounces=8, render_width=W, render_height=H, spp=16)
    out = []
    for builder in (ks.KS_BVH_BUILD_HOST_SBVH, ks.KS_BVH_BUILD_DEVICE):
        sc = ks.Scene(device, cornell, bvh={"builder": builder})
        film = ks.Film(device, W, H)
        ks.small_pt_render(sc, cornell.camera, p, film)
        out.append(film.download())
        film.close()
        sc.close()
    a, wa = out[0]
    for b, wb in out[1:]:
        assert np.array_equal(wa, wb)
        same = np.all(a == b, axis=-1).mean()
        rel = np.abs(a - b).sum() / np.abs(a).sum()
        _log("pixels bit-equal", same, "relative L1 difference", rel)
        assert same >= 0.99 and rel < 1e-3


@pytest.mark.parametrize("builder", BUILDERS)
def test_device_bvh_alpha_tested(device, ks, ora, builder):
    """Alpha-tested triangles keep their flag in the device-built records: the opacity filter gives the
    same hits as over the host SBVH and the oracle."""
    box = textured_box(stochastic_opacity=True)
    sc = ks.Scene(device, box, bvh={"builder": builder})
    nodes, refs = sc.bvh()
    alpha_tris = set(refs[refs[:, 7] != 0, 3].tolist())  # v1.w
    assert alpha_tris == {12, 13, 14, 15}, alpha_tris
    host = ks.Scene(device, box)
    rng = np.random.default_rng(4)
    n = 50000
    o = np.tile(np.float32([0.35, 0.3, 1.5]), (n, 1)) + rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    tgt = np.stack([rng.uniform(0.1, 0.95, n), rng.uniform(0.05, 0.6, n), np.full(n, 0.3)], -1).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = o, d, 100.0
    g = sc.intersect(rays)
    h = host.intersect(rays)
    assert np.array_equal(g[:, 3].view(np.int32), h[:, 3].view(np.int32))
    assert np.array_equal(sc.occluded(rays), host.occluded(rays))
    oo = ora.OracleScene.from_host_scene(box).intersect(rays)
    assert (g[:, 3].view(np.int32) == oo[:, 3].view(np.int32)).mean() > 0.999
    host.close()
    sc.close()


@pytest.mark.parametrize("builder", BUILDERS)
def test_device_bvh_large_scene(device, ks, builder):
    """The 1.3M-triangle sphere scene: structure-level counts, and the same closest hits as the host
    SBVH for a million random rays."""
    hs = ks.procedural_scene("spheres", 1)
    sc = ks.Scene(device, hs, bvh={"builder": builder})
    info = sc.info()
    _log("spheres device build ms", info["bvh_build_ms"], "nodes", info["n_nodes"], "depth", info["max_depth"])
    nodes, refs = sc.bvh()
    assert np.array_equal(np.sort(refs[:, 3]), np.arange(hs.n_tris, dtype=np.uint32))
    host = ks.Scene(device, hs)
    rays = _random_rays(1 << 20, -12.0, 12.0, 9)
    g, h = sc.intersect(rays), host.intersect(rays)
    agree, trel = _compare_hits(g, h)
    _log("agreement with host SBVH", agree, trel)
    assert agree >= 0.99999 and trel == 0.0
    assert np.array_equal(sc.occluded(rays), host.occluded(rays))
    host.close()
    sc.close()


def test_ks_render_device_builder(tmp_path, ks, device, cornell):
    """`bvh_builder = "device"` in a small_pt task renders over the device-built tree: the EXR equals the
    API render over a device-built scene bit for bit."""
    import subprocess
    from pathlib import Path

    from image_files import read_exr_rgb

    tool = Path(__file__).resolve().parents[1] / "keepsake_amd" / "_lib" / "ks_render"
    cfg = tmp_path / "c.toml"
    cfg.write_text(
        f"""
output_dir = "{tmp_path / 'out'}"
[[task]]
type = "small_pt"
task_dir = "cornell"
procedural = "cornell"
bounces = 8
render_width = 64
render_height = 64
spp = 8
bvh_builder = "device"
"""
    )
    r = subprocess.run([str(tool), "-c", str(cfg)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:], flush=True)
    assert r.returncode == 0
    exr, _ = read_exr_rgb(tmp_path / "out" / "cornell" / "small_pt_spp8.exr")
    sc = ks.Scene(device, cornell, bvh={"builder": ks.KS_BVH_BUILD_DEVICE})
    film = ks.Film(device, 64, 64)
    ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(bounces=8, render_width=64, render_height=64, spp=8),
                       film)
    assert np.array_equal(exr, film.composite())
    film.close()
    sc.close()
