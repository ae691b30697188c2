"""GPU parity: the gfx950 kernels (through the C ABI) against the CPU oracle on the same inputs.

Bars: Sobol points bit-exact (uint32 and float); camera rays within 2e-6 absolute; closest-hit
primitive ids equal on >= 99.999% of primary rays (equal-t coplanar ties count as agreement) with t
within 1e-5 relative; BSDF values within 1e-4 relative; light tables equal; renders within Monte Carlo
noise of the oracle at equal spp.
"""

import numpy as np
import pytest

from conftest import rel_mse

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


@pytest.mark.parametrize(
    "W,H,spp,seed",
    [(256, 256, 16, 0), (1920, 1080, 64, 0), (1280, 720, 8, 3), (64, 48, 1, 5), (100, 37, 2, 9), (3840, 2160, 32, 1)],
)
def test_sobol_bit_exact(device, ora, W, H, spp, seed):
    _log("sobol", W, H, spp, seed)
    rng = np.random.default_rng(W + spp)
    n = 512
    px = rng.integers(0, W, n).astype(np.int32)
    py = rng.integers(0, H, n).astype(np.int32)
    s = rng.integers(0, spp, n).astype(np.int32)
    kinds = np.array([2, 1, 2, 1, 2, 1, 1, 2, 1, 2, 1, 1, 2, 1, 2, 1], np.int32)
    gu, gf = device.sobol_stream(W, H, spp, seed, px, py, s, kinds)
    for i in range(n):
        ou, of = ora.sobol(W, H, spp, seed, int(px[i]), int(py[i]), int(s[i]), kinds)
        assert np.array_equal(gu[i], ou), (i, px[i], py[i], s[i])
        assert np.array_equal(gf[i].view(np.uint32), of.view(np.uint32))


def test_camera_rays(device, ora, cornell, spheres_small):
    for hs, W, H, spp in ((cornell, 256, 256, 16), (spheres_small, 1920, 1080, 64), (cornell, 64, 64, 1)):
        rng = np.random.default_rng(1)
        n = 4096
        px = rng.integers(0, W, n).astype(np.int32)
        py = rng.integers(0, H, n).astype(np.int32)
        s = rng.integers(0, spp, n).astype(np.int32)
        g = device.camera_rays(hs.camera, W, H, spp, 0, px, py, s)
        o = ora.camera_rays(hs.camera, W, H, spp, 0, px, py, s)
        assert np.abs(g - o).max() < 2e-6, np.abs(g - o).max()


def _primary_rays(ora, hs, W, H, spp, n, seed=0):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, W, n).astype(np.int32)
    py = rng.integers(0, H, n).astype(np.int32)
    s = rng.integers(0, spp, n).astype(np.int32)
    r6 = ora.camera_rays(hs.camera, W, H, spp, 0, px, py, s)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = r6[:, 0:3]
    rays[:, 4:7] = r6[:, 3:6]
    rays[:, 7] = np.inf
    return rays


def _compare_hits(g, o):
    gp, op = g[:, 3].view(np.int32), o[:, 3].view(np.int32)
    same = gp == op
    both = (gp >= 0) & (op >= 0)
    tie = both & (np.abs(g[:, 0] - o[:, 0]) <= 1e-6 * np.maximum(1.0, np.abs(o[:, 0])))
    agree = (same | tie).mean()
    hit = both & same
    trel = np.abs(g[hit, 0] - o[hit, 0]) / np.maximum(np.abs(o[hit, 0]), 1e-30)
    return agree, (trel.max() if trel.size else 0.0)


def test_closest_hit_primary_cornell(device, ks, ora, cornell):
    sc = ks.Scene(device, cornell)
    osc = ora.OracleScene.from_host_scene(cornell)
    rays = _primary_rays(ora, cornell, 256, 256, 16, 65536)
    g = sc.intersect(rays)
    o = osc.intersect(rays, brute=True)
    agree, trel = _compare_hits(g, o)
    _log("cornell primary agree", agree, "trel", trel)
    assert agree >= 0.99999 and trel <= 1e-5
    sc.close()


def test_closest_hit_primary_spheres(device, ks, ora, spheres_small):
    sc = ks.Scene(device, spheres_small)
    osc = ora.OracleScene.from_host_scene(spheres_small)
    rays = _primary_rays(ora, spheres_small, 1920, 1080, 64, 200000)
    g = sc.intersect(rays)
    o = osc.intersect(rays)
    agree, trel = _compare_hits(g, o)
    _log("spheres primary agree", agree, "trel", trel)
    assert agree >= 0.99999 and trel <= 1e-5
    sc.close()


def test_random_rays_and_occlusion(device, ks, ora, spheres_small):
    sc = ks.Scene(device, spheres_small)
    osc = ora.OracleScene.from_host_scene(spheres_small)
    rng = np.random.default_rng(7)
    n = 100000
    o = rng.uniform(-12, 12, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, d
    rays[:, 7] = rng.uniform(0.5, 30, n).astype(np.float32)
    g = sc.intersect(rays)
    oo = osc.intersect(rays)
    agree, trel = _compare_hits(g, oo)
    assert agree >= 0.9999 and trel <= 1e-5
    occ = sc.occluded(rays)
    assert np.array_equal(occ.astype(bool), oo[:, 3].view(np.int32) >= 0)
    sc.close()


def test_light_table_matches_oracle(device, ks, ora, cornell, spheres_small):
    for hs in (cornell, spheres_small):
        sc = ks.Scene(device, hs)
        a = sc.light_table()
        b = ora.OracleScene.from_host_scene(hs).light_table()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        sc.close()


def _materials():
    m = np.zeros(5, dtype=__import__("keepsake_amd").api.MATERIAL_DTYPE)
    m["basecolor"] = [[0.73, 0.73, 0.73], [0.8, 0.3, 0.2], [0.9, 0.9, 0.9], [0.5, 0.6, 0.7], [0.2, 0.8, 0.4]]
    m["roughness"] = [1.0, 0.3, 0.05, 0.6, 0.4]
    m["metallic"] = [0.0, 1.0, 0.0, 0.4, 0.0]
    m["ior"] = [1.0, 1.5, 1.5, 1.45, 1.33]
    m["specular_r0_mul"] = 0.5
    m["specular_trans"] = [0.0, 0.0, 0.0, 0.3, 0.8]
    m["diffuse_trans"] = [0.0, 0.0, 0.0, 0.2, 0.0]
    m["diffuse_trans_fwd"] = [0.0, 0.0, 0.0, 0.5, 0.0]
    return m


def test_bsdf_eval_sample(device, ora):
    rng = np.random.default_rng(11)
    n = 20000
    for mat in _materials():
        wo = rng.normal(size=(n, 3)).astype(np.float32)
        wo /= np.linalg.norm(wo, axis=1, keepdims=True)
        wo[: n // 2, 2] = np.abs(wo[: n // 2, 2])
        wi = rng.normal(size=(n, 3)).astype(np.float32)
        wi /= np.linalg.norm(wi, axis=1, keepdims=True)
        gf, gp = device.bsdf_eval(mat, wo, wi)
        of, op = ora.bsdf_eval(mat, wo, wi)
        tol = 2e-4 * np.maximum(1.0, np.abs(of))
        assert (np.abs(gf - of) <= tol).mean() > 0.999
        assert (np.abs(gp - op) <= 2e-4 * np.maximum(1.0, np.abs(op))).mean() > 0.999
        u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        gwi, gb, gpd = device.bsdf_sample(mat, wo, u)
        owi, ob, opd = ora.bsdf_sample(mat, wo, u)
        ok = (np.abs(gwi - owi).max(axis=1) < 1e-4) & (np.abs(gpd - opd) <= 1e-3 * np.maximum(1, np.abs(opd)))
        assert ok.mean() > 0.999, ok.mean()


def test_trace_samples_match(device, ks, ora, cornell):
    sc = ks.Scene(device, cornell)
    osc = ora.OracleScene.from_host_scene(cornell)
    W = H = 256
    spp = 16
    rng = np.random.default_rng(2)
    n = 20000
    px = rng.integers(0, W, n).astype(np.int32)
    py = rng.integers(0, H, n).astype(np.int32)
    s = rng.integers(0, spp, n).astype(np.int32)
    kw = dict(bounces=8, render_width=W, render_height=H, spp=spp)
    g = sc.trace_samples(cornell.camera, ks.small_pt_params(**kw), px, py, s)
    o = osc.trace_samples(cornell.camera, ora.params(**kw), px, py, s)
    close = np.all(np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o)), axis=1)
    _log("per-sample agreement", close.mean())
    # the same Sobol points drive both; paths only diverge after float-rounding flips a decision
    assert close.mean() > 0.97
    assert np.array_equal(g[:, 3], o[:, 3])
    sc.close()


def test_render_cornell_matches_oracle(device, ks, ora, cornell):
    W = H = 256
    kw = dict(bounces=8, render_width=W, render_height=H, spp=16, rng_seed=0)
    sc = ks.Scene(device, cornell)
    film = ks.Film(device, W, H)
    st = ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(**kw), film)
    grgb, gw = film.download()
    orgb, ow = ora.OracleScene.from_host_scene(cornell).render(cornell.camera, ora.params(**kw))
    assert np.array_equal(gw, ow)
    gimg = grgb / np.maximum(gw[..., :1], 1)
    oimg = orgb / np.maximum(ow[..., :1], 1)
    e = rel_mse(gimg, oimg)
    _log("cornell relMSE gpu vs oracle", e, "stats", st["iterations"], st["rays_closest"], st["nonfinite"])
    assert e < 1e-3
    assert st["nonfinite"] == 0
    film.close()
    sc.close()
