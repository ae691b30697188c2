"""ks_film_relmse_noalloc (the allocation-free relMSE over a static device scratch) against ks_film_relmse:
two calls in a row on one device, the second on smaller films, each to 1e-12 relative (FP64 sums in a
different order). Run alone or after the other GPU tests: the scratch keeps the last call's contents."""

import pytest

pytestmark = pytest.mark.gpu


def _films(ks, device, sc, cam, W, H, spp_a, spp_b):
    a = ks.Film(device, W, H, backdrop=(0.1, 0.2, 0.3))
    b = ks.Film(device, W, H, backdrop=(0.1, 0.2, 0.3))
    p = dict(bounces=6, render_width=W, render_height=H, include_background=False, backdrop=(0.1, 0.2, 0.3))
    ks.small_pt_render(sc, cam, ks.small_pt_params(spp=spp_a, **p), a)
    ks.small_pt_render(sc, cam, ks.small_pt_params(spp=spp_b, rng_seed=5, **p), b)
    return a, b


def test_noalloc_relmse_matches_relmse(ks, device, cornell):
    sc = ks.Scene(device, cornell)
    cam = dict(cornell.camera)
    # the first call covers more pixels than the scratch has blocks (grid-stride), the second far fewer
    for W, H in ((1024, 640), (40, 24)):
        cam["aspect"] = W / H
        a, b = _films(ks, device, sc, cam, W, H, 2, 16)
        want = a.relmse(b)
        got = a.relmse_noalloc(b)
        print(W, H, "relMSE", want, "no-alloc", got, flush=True)
        assert want > 0 and abs(got - want) <= 1e-12 * want
        assert abs(b.relmse_noalloc(a) - b.relmse(a)) <= 1e-12 * b.relmse(a)
        a.close()
        b.close()
    x = ks.Film(device, 8, 8)
    y = ks.Film(device, 8, 9)
    with pytest.raises(ks.KsError) as e:
        x.relmse_noalloc(y)
    assert e.value.status == ks.KS_ERR_INVALID
    x.close()
    y.close()
    sc.close()
