"""Pins of the oracle's light table, light selection and light sampling (light.cpp:59-134, 136-240,
387-409, 567-701; distrib.cpp:9-41) and of the film composite (render_target.cpp:127-142, 181-192),
CPU only.

- Powers are closed forms of the scene: 2 pi A lum(Le) for a uniform emitter triangle (light.cpp:387-409),
  4 pi I for a point light (:240), L pi r^2 for a directional light (:218-222) and the 1x1 midpoint sum
  2 pi^2 L pi r^2 of SkyLight::power (:136-158) for the ambient sky, r the scene's bounding radius.
- The CDF is the normalised prefix sum of the powers (FP64), the selection is upper_bound over it and its
  frequencies follow the powers.
- Triangle samples lie on the triangle, are uniform in area (chi^2 over the barycentric simplex) and
  carry the solid-angle pdf d^2 / (A |cos|); MeshTriLight::pdf gives the same value for the same
  direction. The ambient sky's pdf is 1 / (2 pi^2 sin theta) in y-up spherical coordinates, uniform in
  (theta, phi).
- The composites are checked on hand-computed film sums.
"""

import numpy as np
import pytest
from scipy import stats

import ks_image_oracle as kio
import ks_oracle as ko
from keepsake_amd.api import LIGHT_DTYPE, MATERIAL_DTYPE

LUM = np.array([0.212671, 0.715160, 0.072169])


def _quad(c, u, v):
    c, u, v = (np.asarray(x, np.float64) for x in (c, u, v))
    return [c, c + u, c + u + v, c + v]


def _light_scene():
    """A floor quad, three emissive quads (different areas, orientations and radiance), an ambient sky,
    a point light and a directional light."""
    quads = [
        (_quad([-5, 0, -5], [10, 0, 0], [0, 0, 10]), None),
        (_quad([-0.5, 3, -0.5], [1, 0, 0], [0, 0, 1]), (17.0, 12.0, 4.0)),
        (_quad([2, 1, -1], [0, 0, 0.5], [0, 2, 0]), (1.0, 5.0, 0.5)),
        (_quad([-3, 0.5, 2], [1.5, 0.3, 0], [0, 1.2, 0.7]), (3.0, 3.0, 3.0)),
    ]
    verts, idx, tri_geom = [], [], []
    for g, (q, _) in enumerate(quads):
        b = len(verts)
        verts += q
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        tri_geom += [g, g]
    mats = np.zeros(len(quads), MATERIAL_DTYPE)
    mats["basecolor"] = 0.5
    mats["ior"] = 1.5
    for g, (_, le) in enumerate(quads):
        if le is not None:
            mats[g]["emission"] = le
            mats[g]["has_emission"] = 1
    lights = np.zeros(3, LIGHT_DTYPE)
    lights[0]["type"], lights[0]["color"] = 0, (0.5, 0.6, 0.7)
    lights[1]["type"], lights[1]["color"], lights[1]["vec"] = 2, (30.0, 20.0, 10.0), (0.0, 6.0, 0.0)
    lights[2]["type"], lights[2]["color"], lights[2]["vec"] = 1, (2.0, 2.0, 2.0), (0.3, 1.0, 0.2)
    verts = np.asarray(verts, np.float32)
    idx = np.asarray(idx, np.uint32)
    sc = ko.OracleScene(verts, idx, np.asarray(tri_geom, np.uint32), np.arange(len(quads), dtype=np.uint32), mats,
                        lights)
    return sc, verts, idx, mats, lights


@pytest.fixture(scope="module")
def light_scene():
    return _light_scene()


def _tri_area(verts, idx, t):
    a, b, c = (verts[idx[t, k]].astype(np.float64) for k in range(3))
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a))


def test_light_table_powers_and_cdf(light_scene):
    sc, verts, idx, mats, lights = light_scene
    prim, power, cdf = sc.light_table()
    # flat order: the non-mesh lights as given, then every emissive triangle in geometry order
    assert list(prim[:3]) == [-1, -1, -1]
    assert list(prim[3:]) == [2, 3, 4, 5, 6, 7]
    lo, hi = verts.min(0).astype(np.float64), verts.max(0).astype(np.float64)
    r = 0.5 * np.linalg.norm(hi - lo)
    want = [np.mean(lights[0]["color"]) * 2 * np.pi**2 * np.pi * r**2,
            np.mean(lights[1]["color"]) * 4 * np.pi,
            np.mean(lights[2]["color"]) * np.pi * r**2]
    for t in prim[3:]:
        le = np.asarray(mats[t // 2]["emission"], np.float64)
        want.append(2 * np.pi * _tri_area(verts, idx, t) * float(LUM @ le))
    want = np.array(want)
    assert np.allclose(power, want, rtol=2e-6), (power, want)
    c = np.concatenate([[0.0], np.cumsum(want)]) / want.sum()
    assert cdf[0] == 0.0 and cdf[-1] == 1.0
    assert np.allclose(cdf, c, rtol=1e-6, atol=1e-7)


def test_light_selection_follows_powers(light_scene):
    sc, *_ = light_scene
    _, power, cdf = sc.light_table()
    u = ((np.arange(1_000_000) + 0.5) / 1_000_000).astype(np.float32)
    u = np.concatenate([u, np.float32([0.0, np.nextafter(np.float32(1), np.float32(0)), 1.0])])
    i, prob = sc.light_select(u)
    # std::upper_bound over the CDF, u clamped below one (distrib.cpp:33-41)
    uc = np.minimum(u, np.nextafter(np.float32(1), np.float32(0)))
    assert np.array_equal(i, np.searchsorted(cdf, uc, side="right") - 1)
    assert np.allclose(prob, np.diff(cdf)[i], rtol=1e-6)
    freq = np.bincount(i[:1_000_000], minlength=len(power)) / 1e6
    assert np.allclose(freq, power / power.sum(), atol=2e-6)


def test_triangle_samples_uniform_with_solid_angle_pdf(light_scene):
    sc, verts, idx, mats, _ = light_scene
    prim, _, _ = sc.light_table()
    rng = np.random.default_rng(3)
    n = 100_000
    for li in range(3, len(prim)):
        t = prim[li]
        a, b, c = (verts[idx[t, k]].astype(np.float64) for k in range(3))
        area = _tri_area(verts, idx, t)
        nrm = np.cross(b - a, c - b)
        nrm /= np.linalg.norm(nrm)
        p = (rng.uniform(-4, 4, (n, 3)) * [1, 0, 1] + [0, 0.2, 0]).astype(np.float32)
        u = rng.uniform(0, 1, (n, 2)).astype(np.float32)
        wi, dist, pdf, le_over = sc.light_sample(li, p, u)
        ok = pdf > 0
        assert ok.mean() > 0.999
        wi, dist, pdf, le_over, p = wi[ok], dist[ok], pdf[ok], le_over[ok], p[ok]
        d = dist.astype(np.float64) + 1e-4  # the sample shortens the shadow ray by 1e-4 (light.cpp:600-601)
        x = p.astype(np.float64) + wi.astype(np.float64) * d[:, None]
        # on the triangle: barycentrics from FP64 areas
        m = np.stack([b - a, c - a], 0)
        bary = np.linalg.lstsq(m.T, (x - a).T, rcond=None)[0].T
        assert np.all(np.abs((x - a - bary @ m)) < 1e-4)
        assert bary.min() > -1e-4 and (bary.sum(1)).max() < 1 + 1e-4
        # solid-angle pdf d^2 / (A |cos|) and Le / pdf
        cos = np.abs(wi.astype(np.float64) @ nrm)
        want = d**2 / (area * cos)
        good = cos > 1e-3
        assert np.allclose(pdf[good], want[good], rtol=2e-4)
        le = np.asarray(mats[t // 2]["emission"], np.float64)
        assert np.allclose(le_over[good], le[None] / pdf[good, None], rtol=1e-5)
        # MeshTriLight::pdf of the same direction at the hit distance
        p2 = sc.light_pdf(li, wi[good], d[good].astype(np.float32))
        assert np.allclose(p2, pdf[good], rtol=2e-4)
        # uniform in area: chi^2 over an 8 x 8 grid of the barycentric simplex (diagonal cells half area)
        k = 8
        ci = np.clip((bary * k).astype(int), 0, k - 1)
        obs = np.bincount(ci[:, 0] * k + ci[:, 1], minlength=k * k).astype(np.float64)
        ii, jj = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
        cell = np.where(ii + jj < k - 1, 1.0, np.where(ii + jj == k - 1, 0.5, 0.0)).ravel()
        exp = cell / cell.sum() * len(bary)
        keep = exp > 0
        chi2 = (((obs - exp) ** 2)[keep] / exp[keep]).sum()
        assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4 and obs[~keep].sum() <= 2 * n * 1e-4


def test_triangle_sample_mapping(light_scene):
    """The square-to-triangle map is Heitz's low-distortion map (Heitz 2019, "A Low-Distortion Map Between
    Triangle and Square"; sample_triangle_bary, rng.h:362-373): (b0, b1) = (u0 / 2, u1 - u0 / 2) if
    u0 < u1, else (u0 - u1 / 2, u1 / 2), with the point (1 - b0 - b1) v0 + b0 v1 + b1 v2 in Embree's
    barycentric convention (rtcInterpolate0 at (u, v) = (b0, b1))."""
    sc, verts, idx, _, _ = light_scene
    prim, _, _ = sc.light_table()
    li = 4
    t = prim[li]
    v0, v1, v2 = (verts[idx[t, k]].astype(np.float64) for k in range(3))
    u = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0], [1.0, 1.0], [0.25, 0.75], [0.75, 0.25], [0.5, 0.5]])
    b0 = np.where(u[:, 0] < u[:, 1], u[:, 0] / 2, u[:, 0] - u[:, 1] / 2)
    b1 = np.where(u[:, 0] < u[:, 1], u[:, 1] - u[:, 0] / 2, u[:, 1] / 2)
    want = v0 * (1 - b0 - b1)[:, None] + v1 * b0[:, None] + v2 * b1[:, None]
    p = np.array([0.0, 0.2, 0.0], np.float32)
    wi, dist, _, _ = sc.light_sample(li, p, u.astype(np.float32))
    x = p.astype(np.float64) + wi.astype(np.float64) * (dist.astype(np.float64) + 1e-4)[:, None]
    assert np.allclose(x, want, atol=2e-5), (x, want)


def test_non_mesh_light_samples(light_scene):
    sc, verts, idx, mats, lights = light_scene
    rng = np.random.default_rng(4)
    n = 200_000
    p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    u = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    # ambient sky: uniform in (theta, phi) about +y, pdf 1 / (2 pi^2 sin theta), Le / pdf
    wi, dist, pdf, le_over = sc.light_sample(0, p, u)
    ok = pdf > 0
    assert ok.mean() > 0.9999 and np.all(np.isinf(dist[ok]))
    y = wi[ok, 1].astype(np.float64)
    sin_t = np.hypot(wi[ok, 0].astype(np.float64), wi[ok, 2].astype(np.float64))  # exact near the poles
    assert np.allclose(pdf[ok], 1 / (2 * np.pi**2 * sin_t), rtol=1e-4)
    assert np.allclose(le_over[ok], np.asarray(lights[0]["color"])[None] * (2 * np.pi**2 * sin_t)[:, None], rtol=1e-4)
    theta = np.arccos(np.clip(y, -1, 1)) / np.pi
    phi = (np.arctan2(wi[ok, 2], wi[ok, 0]) % (2 * np.pi)) / (2 * np.pi)
    assert stats.kstest(theta, "uniform").pvalue > 1e-4 and stats.kstest(phi, "uniform").pvalue > 1e-4
    # point light: towards the position, pdf 1, I / d^2
    wi, dist, pdf, le_over = sc.light_sample(1, p, u)
    v = np.asarray(lights[1]["vec"], np.float64) - p.astype(np.float64)
    d = np.linalg.norm(v, axis=1)
    assert np.allclose(wi, v / d[:, None], atol=1e-6) and np.allclose(dist, d, rtol=1e-6) and np.all(pdf == 1)
    assert np.allclose(le_over, np.asarray(lights[1]["color"])[None] / d[:, None] ** 2, rtol=1e-5)
    # directional light: the normalised direction, infinitely far, radiance unchanged
    wi, dist, pdf, le_over = sc.light_sample(2, p[:100], u[:100])
    dd = np.asarray(lights[2]["vec"], np.float64)
    assert np.allclose(wi, dd / np.linalg.norm(dd), atol=1e-6) and np.all(np.isinf(dist)) and np.all(pdf == 1)
    assert np.allclose(le_over, np.asarray(lights[2]["color"])[None])


def test_composites_on_hand_built_sums():
    """RenderTarget2 composites (render_target.cpp:127-142, 181-192) on sums whose result is known."""
    sums = np.array([[3.0, 6.0, 9.0], [0.0, 0.0, 0.0], [2.0, 2.0, 2.0], [0.0, 0.0, 0.0], [5.0, 1.0, 0.5]])
    w = np.array([[3.0, 1.0], [0.0, 4.0], [2.0, 0.0], [0.0, 0.0], [1.0, 1.0]])
    bd = np.array([1.0, 0.5, 0.25])
    exr = kio.composite_exr(sums[None], w[None], bd)[0]
    # lerp(backdrop, sum / w_hit, w_hit / (w_hit + w_miss)), backdrop where nothing hit, 0 where empty
    want = np.array([[0.25 * 1 + 0.75 * 1, 0.25 * 0.5 + 0.75 * 2, 0.25 * 0.25 + 0.75 * 3],
                     [1.0, 0.5, 0.25], [1.0, 1.0, 1.0], [1.0, 0.5, 0.25], [3.0, 0.75, 0.375]])
    assert np.allclose(exr, want, rtol=1e-7), exr
    # the PNG form (sum + w_miss backdrop) / (w_hit + w_miss) agrees wherever w > 0
    png_lin = (sums + w[:, 1:] * bd) / np.maximum(w.sum(1, keepdims=True), 1e-300)
    assert np.allclose(png_lin[[0, 1, 2, 4]], want[[0, 1, 2, 4]], rtol=1e-7)
    b = kio.composite_png(sums[None, [0, 4]], w[None, [0, 4]], bd)[0]
    assert np.array_equal(b, kio.srgb_bytes(want[[0, 4]].astype(np.float32)))
