"""The measurement entry points: ks_trace_capture returns exactly the rays a render's paths issue, and
ks_scene_traversal_counts traverses them in the reference order (cuda/sbvh.cu:512-603) to the same hits
as the oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _paths(W, H, spp, n, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, W, n).astype(np.int32), rng.integers(0, H, n).astype(np.int32),
            rng.integers(0, spp, n).astype(np.int32))


@pytest.mark.parametrize("scene", ["cornell", "spheres_small"])
def test_captured_rays_and_reference_order_counts(ks, device, ora, scene):
    hs = ks.procedural_scene(scene, 1 if scene != "cornell" else 0)
    sc = ks.Scene(device, hs)
    W, H, spp = 128, 96, 16
    kw = dict(bounces=8, render_width=W, render_height=H, spp=spp)
    cam = dict(hs.camera)
    cam["aspect"] = W / H
    n = 4096
    px, py, s = _paths(W, H, spp, n, 3)
    rc, rs = sc.capture_rays(cam, ks.small_pt_params(**kw), px, py, s, max_rays=n * 10)
    print(scene, "captured", len(rc), "closest and", len(rs), "shadow rays", flush=True)
    assert len(rc) >= n and len(rs) > 0
    # the first n closest rays are the primary rays of the listed paths, in order
    prim = device.camera_rays(cam, W, H, spp, 0, px, py, s)
    assert np.array_equal(rc[:n, 0:3], prim[:, 0:3]) and np.array_equal(rc[:n, 4:7], prim[:, 3:6])
    # every captured ray is a valid query: unit direction, tmin 0 <= tmax
    assert np.allclose(np.linalg.norm(rc[:, 4:7], axis=1), 1.0, atol=1e-5)
    assert np.all(rc[:, 3] >= 0) and np.all(rc[:, 7] >= rc[:, 3])
    # the reference-order traversal finds the oracle's closest hits and occlusion
    nodes, tris, t = sc.traversal_counts(rc, with_t=True)
    osc = ora.OracleScene.from_host_scene(hs)
    o = osc.intersect(rc)
    ot = np.where(o[:, 3].view(np.int32) >= 0, o[:, 0], -1.0)
    same = np.isclose(t, ot, rtol=1e-6, atol=0) | ((t < 0) & (ot < 0))
    print(scene, "closest agree", same.mean(), "nodes/ray", nodes / len(rc), "tris/ray", tris / len(rc), flush=True)
    assert same.mean() > 0.9999
    ns, ts, occ = sc.traversal_counts(rs, any_hit=True, with_t=True)
    oocc = osc.occluded(rs)
    assert ((occ > 0) == (oocc == 1)).mean() > 0.9999
    assert ns >= len(rs) and nodes >= len(rc) and tris > 0
    osc.close()
    sc.close()


def test_counts_of_a_missing_ray(ks, device, cornell):
    """A ray that misses the scene's bound costs exactly the root's fetch."""
    sc = ks.Scene(device, cornell)
    r = np.array([[5, 5, 5, 0, 1, 0, 0, np.inf]], np.float32)
    assert sc.traversal_counts(r) == (1, 0)
    assert sc.traversal_counts(r, any_hit=True) == (1, 0)
    sc.close()


def test_counts_need_the_host_sbvh(ks, device, cornell):
    sc = ks.Scene(device, cornell, bvh={"builder": 1})
    with pytest.raises(ks.KsError) as e:
        sc.traversal_counts(np.zeros((1, 8), np.float32))
    assert e.value.status == ks.KS_ERR_UNSUPPORTED
    sc.close()
