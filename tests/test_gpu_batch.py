"""Batched renders (ks_small_pt_render_batch): many views of one scene in one wavefront. Every view's film
must equal, bit for bit, the film of that view rendered alone: the same Sobol samples per pixel, folded
in the same order. Covered: mixed resolutions, spp (including non powers of two), seeds, a crop and a
moved camera; graph replay and eager launches; the autotuner and a fixed queue."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _views(ks, cornell):
    cam2 = dict(cornell.camera)
    cam2["pos"] = (0.2, 0.6, 2.0)
    specs = [
        (cornell.camera, dict(render_width=64, render_height=64, spp=4, rng_seed=0)),
        (cornell.camera, dict(render_width=128, render_height=72, spp=16, rng_seed=3)),
        (cam2, dict(render_width=96, render_height=96, spp=5, rng_seed=1)),
        (cornell.camera, dict(render_width=200, render_height=120, spp=8, rng_seed=2, crop_start_x=20,
                              crop_start_y=10, crop_width=100, crop_height=50)),
        (cam2, dict(render_width=32, render_height=48, spp=64, rng_seed=7)),
    ]
    out = []
    for cam, kw in specs:
        p = ks.small_pt_params(bounces=12, spp_prog_interval=kw["spp"], **kw)
        out.append((cam, p, kw))
    return out


@pytest.mark.parametrize("mode", ["graph_autotune", "eager_fixed"])
def test_batch_equals_single_views(ks, device, cornell, mode):
    sc = ks.Scene(device, cornell)
    views = _views(ks, cornell)
    films = [ks.Film(device, kw["render_width"], kw["render_height"]) for _, _, kw in views]
    if mode == "graph_autotune":
        st = ks.small_pt_render_batch(sc, [(c, p, f) for (c, p, _), f in zip(views, films)])
    else:
        st = ks.small_pt_render_batch(sc, [(c, p, f) for (c, p, _), f in zip(views, films)], queue_size=1 << 14,
                                      trace_blocks_per_cu=4, graph=False)
    print(mode, "queue", st["queue_size"], "blocks/CU", st["trace_blocks_per_cu"], "paths", st["paths"],
          "iterations", st["iterations"], flush=True)
    assert st["paths"] == sum(kw["spp"] * (kw.get("crop_width") or kw["render_width"]) *
                              (kw.get("crop_height") or kw["render_height"]) for _, _, kw in views)
    for (cam, p, kw), f in zip(views, films):
        alone = ks.Film(device, kw["render_width"], kw["render_height"])
        ks.small_pt_render(sc, cam, p, alone)
        a, wa = f.download()
        b, wb = alone.download()
        assert np.array_equal(wa, wb), kw
        assert np.array_equal(a, b), kw
        alone.close()
    for f in films:
        f.close()
    sc.close()


@pytest.mark.parametrize("graph", [True, False])
def test_views_split_across_chunks(ks, device, cornell, graph):
    """A small chunk budget (50 000 items) splits the batch into four chunks, and the spp of the first
    three views across chunk boundaries (segments starting at s0 > 0, one sentinel per chunk): every film
    still equals its single render bit for bit. The autotuner's probe runs on the first chunk's partial
    segment list."""
    sc = ks.Scene(device, cornell)
    cam2 = dict(cornell.camera)
    cam2["pos"] = (0.3, 0.4, 2.2)
    specs = [(cornell.camera, dict(render_width=64, render_height=64, spp=16, rng_seed=1)),
             (cam2, dict(render_width=32, render_height=32, spp=64, rng_seed=2)),
             (cornell.camera, dict(render_width=48, render_height=40, spp=9, rng_seed=3))]
    views = [(cam, ks.small_pt_params(bounces=12, spp_prog_interval=kw["spp"], **kw), kw) for cam, kw in specs]
    films = [ks.Film(device, kw["render_width"], kw["render_height"]) for _, _, kw in views]
    kwargs = dict(max_chunk_items=50_000, graph=graph)
    if not graph:
        kwargs.update(queue_size=1 << 13, trace_blocks_per_cu=4)
    st = ks.small_pt_render_batch(sc, [(c, p, f) for (c, p, _), f in zip(views, films)], **kwargs)
    print("chunks", st["batch_chunks"], "queue", st["queue_size"], flush=True)
    assert st["batch_chunks"] == 4
    for (cam, p, kw), f in zip(views, films):
        alone = ks.Film(device, kw["render_width"], kw["render_height"])
        ks.small_pt_render(sc, cam, p, alone)
        a, wa = f.download()
        b, wb = alone.download()
        assert np.array_equal(wa, wb) and np.array_equal(a, b), kw
        alone.close()
    for f in films:
        f.close()
    sc.close()
