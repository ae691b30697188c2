"""Pixel filters on the GPU (SmallPTInput::pixel_filter; render_target.h:70-184): the gaussian and
Blackman-Harris sample tables and a wider box give the oracle's camera jitter, hence the oracle's
per-sample radiance and image."""

import numpy as np
import pytest

from conftest import rel_mse

pytestmark = pytest.mark.gpu

FILTERS = {
    "gaussian": dict(pixel_filter=1, filter_width=3.0, filter_sigma=0.5),
    "blackman_harris": dict(pixel_filter=2, filter_width=3.0),
    "box_2": dict(pixel_filter=0, filter_width=2.0),
    "gaussian_narrow": dict(pixel_filter=1, filter_width=2.0, filter_sigma=0.3),
}


@pytest.mark.parametrize("name", list(FILTERS))
def test_filtered_samples_and_image_match_oracle(device, ks, ora, cornell, name):
    f = FILTERS[name]
    sc = ks.Scene(device, cornell)
    osc = ora.OracleScene.from_host_scene(cornell)
    W = H = 128
    spp = 16
    kw = dict(bounces=8, render_width=W, render_height=H, spp=spp, **f)
    rng = np.random.default_rng(5)
    n = 20000
    px = rng.integers(0, W, n).astype(np.int32)
    py = rng.integers(0, H, n).astype(np.int32)
    s = rng.integers(0, spp, n).astype(np.int32)
    g = sc.trace_samples(cornell.camera, ks.small_pt_params(**kw), px, py, s)
    o = osc.trace_samples(cornell.camera, ora.params(**kw), px, py, s)
    close = np.all(np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o)), axis=1)
    print(name, "per-sample agreement", close.mean(), flush=True)
    assert close.mean() > 0.999
    assert np.array_equal(g[:, 3], o[:, 3])
    film = ks.Film(device, W, H)
    ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(**kw), film)
    grgb, gw = film.download()
    orgb, ow = osc.render(cornell.camera, ora.params(**kw))
    assert np.array_equal(gw, ow)
    e = rel_mse(grgb / np.maximum(gw[..., :1], 1), orgb / np.maximum(ow[..., :1], 1))
    print(name, "relMSE vs oracle", e, flush=True)
    # one path that parts from the oracle's (a float flip) moves this 128 x 128 x 16 image by up to
    # (clamp_indirect / spp)^2 / 0.01 / (3 W H) = 8e-4; measured 1e-8 .. 1.4e-4
    assert e < 5e-4
    film.close()
    sc.close()


def test_filter_changes_the_image(device, ks, cornell):
    """The filter is honoured: a Blackman-Harris film differs from the box film of the same samples."""
    sc = ks.Scene(device, cornell)
    imgs = []
    for f in ({}, FILTERS["blackman_harris"]):
        film = ks.Film(device, 64, 64)
        ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(bounces=4, render_width=64, render_height=64, spp=4, **f), film)
        rgb, w = film.download()
        imgs.append(rgb / np.maximum(w[..., :1], 1))
        film.close()
    assert rel_mse(imgs[0], imgs[1]) > 1e-3
    sc.close()


def test_bad_filter_is_rejected(device, ks, cornell):
    sc = ks.Scene(device, cornell)
    film = ks.Film(device, 8, 8)
    for f in (dict(pixel_filter=7), dict(pixel_filter=1, filter_width=-1.0)):
        with pytest.raises(ks.KsError) as e:
            ks.small_pt_render(sc, cornell.camera, ks.small_pt_params(render_width=8, render_height=8, spp=4, **f), film)
        assert e.value.status == ks.KS_ERR_INVALID
    film.close()
    sc.close()
