"""Pins of the oracle's ray/triangle test (the watertight test of cuda/ray_tri.cuh:46-97 with its FP64
edge fallback) and of its closest-hit / any-hit queries, CPU only.

Watertightness: rays from inside closed meshes never escape, and rays aimed exactly at shared vertices
and edges of a triangle fan never fall through the triangle test (brute force over all triangles). Hit
distances equal the analytic distance to the hit plane (FP64). The interval rule is strict at both ends
(t > tmin, t < tmax, `ray_tri.cuh:84-88`).

The BVH query uses the reference's slab test (cuda/aabb.cuh:172-251), which has no rounding margin: a
ray aimed exactly at a node box's boundary (a shared vertex on a flat box's corner) can be culled by
one ulp. On such adversarial rays the BVH query may lose a few hits, and every hit it reports is the
brute-force hit. Random rays lose none.
"""

import numpy as np
import pytest

import ks_oracle as ko
from keepsake_amd.api import LIGHT_DTYPE, MATERIAL_DTYPE


def _scene(verts, tris):
    verts = np.asarray(verts, np.float32)
    tris = np.asarray(tris, np.uint32)
    mats = np.zeros(1, MATERIAL_DTYPE)
    mats["basecolor"] = 0.5
    mats["ior"] = 1.5
    return ko.OracleScene(verts, tris, np.zeros(len(tris), np.uint32), np.zeros(1, np.uint32), mats,
                          np.zeros(0, LIGHT_DTYPE))


def _cube():
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    idx = lambda x, y, z: (x > 0) * 4 + (y > 0) * 2 + (z > 0)  # noqa: E731
    tris = []
    for axis in range(3):
        for s in (-1, 1):
            # the face's four corners in cyclic order
            corners = []
            for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = [0, 0, 0]
                p[axis] = s
                p[(axis + 1) % 3] = a
                p[(axis + 2) % 3] = b
                corners.append(idx(*p))
            tris += [[corners[0], corners[1], corners[2]], [corners[0], corners[2], corners[3]]]
    return v, np.array(tris, np.uint32)


def _icosphere(level):
    t = (1 + 5**0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    v = [list(np.array(p) / np.linalg.norm(p)) for p in v]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = (np.array(v[a]) + np.array(v[b])) / 2
                v.append(list(m / np.linalg.norm(m)))
                cache[k] = len(v) - 1
            return cache[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(v, np.float32), np.array(f, np.uint32)


def _rays(o, d, tmin=0.0, tmax=np.inf):
    n = len(d)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def _prim(hits):
    return hits[:, 3].view(np.int32)


def _check_watertight(sc, rays, min_bvh_frac):
    """Brute force never misses; the BVH misses at most 1 - min_bvh_frac and agrees with brute force on
    every ray it hits (to an ulp or two: through a shared edge both triangles are hit, at distances
    computed from different vertices, and a culled box can leave the other one)."""
    hb = sc.intersect(rays, brute=True)
    assert np.all(_prim(hb) >= 0), f"{np.sum(_prim(hb) < 0)} of {len(rays)} rays fell through the triangle test"
    h = sc.intersect(rays)
    hit = _prim(h) >= 0
    assert hit.mean() >= min_bvh_frac, hit.mean()
    assert np.allclose(h[hit, 0], hb[hit, 0], rtol=1e-6, atol=0)
    return hb


def test_rays_from_inside_a_cube_never_escape():
    v, tris = _cube()
    sc = _scene(v, tris)
    rng = np.random.default_rng(1)
    n = 60000
    o = rng.uniform(-0.95, 0.95, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a third of the rays aim exactly at a cube corner or an edge point of the boundary
    k = n // 3
    target = np.sign(rng.normal(size=(k, 3)))
    edge = rng.integers(0, 3, k)
    target[np.arange(k)[: k // 2], edge[: k // 2]] = rng.uniform(-1, 1, k // 2)
    d[:k] = target - o[:k]
    d[:k] /= np.linalg.norm(d[:k], axis=1, keepdims=True)
    d = d.astype(np.float32)
    h = _check_watertight(sc, _rays(o, d), 0.999)
    assert np.all(_prim(sc.intersect(_rays(o[k:], d[k:]))) >= 0), "random rays escaped the BVH"
    # t equals the distance to the first face plane along the ray (FP64)
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    with np.errstate(divide="ignore"):
        tf = np.where(d64 != 0, (np.sign(d64) - o64) / d64, np.inf)
    t_exact = tf.min(1)
    assert np.allclose(h[:, 0], t_exact, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("level", [2, 4])
def test_rays_through_icosphere_vertices_and_edges_never_escape(level):
    v, tris = _icosphere(level)
    sc = _scene(v, tris)
    rng = np.random.default_rng(level)
    # every vertex and every edge midpoint (in float) as a target, from random interior origins
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    mids = (v[e[:, 0]] + v[e[:, 1]]) * np.float32(0.5)
    quarter = v[e[:, 0]] * np.float32(0.25) + v[e[:, 1]] * np.float32(0.75)
    targets = np.concatenate([v, mids, quarter])
    o = rng.uniform(-0.3, 0.3, (len(targets), 3)).astype(np.float32)
    d = (targets - o).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    h = _check_watertight(sc, _rays(o, d), 0.998)
    assert sc.occluded(_rays(o, d)).mean() >= 0.998
    # the hit lies on its triangle's plane
    p = _prim(h)
    a, b, c = (v[tris[p, k]].astype(np.float64) for k in range(3))
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    t_plane = np.sum((a - o) * nrm, 1) / np.sum(d.astype(np.float64) * nrm, 1)
    assert np.allclose(h[:, 0], t_plane, rtol=1e-5, atol=1e-6)


def test_fan_shared_vertex_and_edges():
    """16 triangles around a shared centre on the plane z = 0.25; rays from above aimed at the centre and
    at float points of the shared edges hit, with t = (z0 - oz) / dz."""
    n = 16
    ang = np.arange(n) * 2 * np.pi / n
    ring = np.stack([np.cos(ang), np.sin(ang), np.full(n, 0.25)], 1)
    v = np.concatenate([[[0.0, 0.0, 0.25]], ring]).astype(np.float32)
    tris = np.array([[0, 1 + i, 1 + (i + 1) % n] for i in range(n)], np.uint32)
    sc = _scene(v, tris)
    rng = np.random.default_rng(5)
    m = 20000
    k = rng.integers(0, n, m)
    s = rng.uniform(0, 0.999, m).astype(np.float32)
    targets = v[0] * (1 - s[:, None]) + v[1 + k] * s[:, None]  # on the shared edge centre -> ring vertex k
    targets[: m // 10] = v[0]  # exactly the shared vertex
    o = (targets + rng.uniform(-2, 2, (m, 3)) * np.array([1, 1, 0]) + np.array([0, 0, 1.5])).astype(np.float32)
    d = (targets - o).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    h = _check_watertight(sc, _rays(o, d), 0.98)
    t = (0.25 - o[:, 2].astype(np.float64)) / d[:, 2].astype(np.float64)
    assert np.allclose(h[:, 0], t, rtol=1e-5)
    # barycentrics of the hit reproduce the hit point (u, v weight vertices 1 and 2 of the triangle)
    p = _prim(h)
    a, b, c = (v[tris[p, j]].astype(np.float64) for j in range(3))
    u, w = h[:, 1].astype(np.float64), h[:, 2].astype(np.float64)
    x = a * (1 - u - w)[:, None] + b * u[:, None] + c * w[:, None]
    y = o.astype(np.float64) + d.astype(np.float64) * h[:, :1].astype(np.float64)
    assert np.allclose(x, y, atol=2e-5)


@pytest.mark.parametrize("winding", [[0, 1, 2], [0, 2, 1]])
def test_interval_is_strict_at_both_ends(winding):
    """Both signs of the determinant (both windings) reject t == tmin and t == tmax."""
    v = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    sc = _scene(v, [winding])
    o = np.array([[0.1, 0.0, 1.0]] * 4, np.float32)
    d = np.array([[0.0, 0.0, -1.0]] * 4, np.float32)
    one = np.float32(1.0)
    r = _rays(o, d)
    r[:, 3] = [0.0, 0.0, one, np.nextafter(one, np.float32(0))]
    r[:, 7] = [one, np.nextafter(one, np.float32(2)), np.inf, np.inf]
    for brute in (False, True):  # the BVH's slab test alone would already cull t == tmax
        h = sc.intersect(r, brute=brute)
        assert list(_prim(h)) == [-1, 0, -1, 0]
        assert h[1, 0] == 1.0 and h[3, 0] == 1.0
    assert list(sc.occluded(r)) == [0, 1, 0, 1]
    # a ray along the plane, and one pointing away, miss
    r2 = _rays(np.array([[-2, 0, 0], [0.1, 0, 1]], np.float32), np.array([[1, 0, 0], [0, 0, 1]], np.float32))
    assert list(_prim(sc.intersect(r2))) == [-1, -1]


def test_closest_of_stacked_triangles():
    """40 parallel triangles at z = k / 8: the closest hit from above is the top one, from below the
    bottom one, and a ray starting between layers (tmin = 0) hits the next layer in its direction."""
    base = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    v = np.concatenate([base + np.array([0, 0, k / 8], np.float32) for k in range(40)])
    tris = np.arange(120, dtype=np.uint32).reshape(40, 3)
    sc = _scene(v, tris)
    o = np.array([[0.0, 0.0, 10.0], [0.0, 0.0, -1.0], [0.0, 0.0, 2.05], [0.0, 0.0, 2.05]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, 1], [0, 0, -1]], np.float32)
    h = sc.intersect(_rays(o, d))
    assert list(_prim(h)) == [39, 0, 17, 16]
    assert np.allclose(h[:, 0], [10 - 39 / 8, 1.0, 17 / 8 - 2.05, 2.05 - 2.0], rtol=1e-6)


def test_edge_function_fp64_fallback():
    """A ray 2^-24 (in edge-function units) off the shared edge BC of two triangles: the float edge
    function rounds to exactly zero, and the FP64 fallback (ray_tri.cuh:61-66) decides the side. Exactly
    one triangle is hit, the one that contains the point in exact (rational) arithmetic."""
    from fractions import Fraction as Fr

    e12, e11 = 2.0**-12, 2.0**-11
    A, A2 = (1.0, -1.0), (-1.0, 1.0)
    B, C = (1 + e12, 1.0), (-(1 + e11), -(1 + e12))
    v = np.array([[*A, 0], [*B, 0], [*C, 0], [*A2, 0]], np.float32)
    tris = [[0, 1, 2], [3, 2, 1]]
    sc = _scene(v, tris)
    # the edge function of BC at the origin: float cancels to 0, exact is -2^-24
    assert np.float32(B[0]) * np.float32(C[1]) - np.float32(C[0]) * np.float32(B[1]) == 0.0
    assert Fr(B[0]) * Fr(C[1]) - Fr(C[0]) * Fr(B[1]) == Fr(-1, 2**24)

    def inside(a, b, c):  # origin inside triangle (a, b, c) by exact edge functions
        e = [Fr(p[0]) * Fr(q[1]) - Fr(q[0]) * Fr(p[1]) for p, q in ((b, c), (c, a), (a, b))]
        return all(x >= 0 for x in e) or all(x <= 0 for x in e)

    want = [inside(A, B, C), inside(A2, C, B)]
    assert want.count(True) == 1
    r = _rays(np.array([[0.0, 0.0, 1.0]], np.float32), np.array([[0.0, 0.0, -1.0]], np.float32))
    for brute in (False, True):
        assert _prim(sc.intersect(r, brute=brute))[0] == want.index(True)
