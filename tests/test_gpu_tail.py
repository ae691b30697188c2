"""The tail kernel (ks_device_set_tail): once every work item of a render has started and few paths
remain, k_tail carries each remaining path to its end in one thread. Its films must equal the films of the
wavefront alone (tail off) bit for bit: the same traversal hits, the same shading code, and each NEE
contribution added with the same single addition before the next vertex."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _render(ks, device, sc, cam, p, W, H, tail):
    device.set_tail(tail)
    f = ks.Film(device, W, H)
    st = ks.small_pt_render(sc, cam, p, f)
    out = f.download()
    f.close()
    return out, st


@pytest.mark.parametrize("name,seed,bounces,extra", [("cornell", 0, 8, dict()), ("spheres_small", 1, 12, dict()),
                                                     ("room", 3, 16, dict(mis_mode=1, clamp_indirect=0.0))])
def test_tail_equals_wavefront(ks, device, name, seed, bounces, extra):
    hs = ks.procedural_scene(name, seed)
    sc = ks.Scene(device, hs)
    W, H = 160, 96
    cam = dict(hs.camera)
    cam["aspect"] = W / H
    p = ks.small_pt_params(bounces=bounces, render_width=W, render_height=H, spp=16, spp_prog_interval=16, **extra)
    (a, wa), sa = _render(ks, device, sc, cam, p, W, H, 0)
    # every remaining path once all work items have started (the control block is read after every
    # iteration near the end of generation, so paths are still in flight)
    (b, wb), sb = _render(ks, device, sc, cam, p, W, H, 1 << 30)
    device.set_tail(0)  # the default
    print(name, "tail rays", sb["rays_tail"], "wavefront rays", sa["rays_closest"] + sa["rays_shadow"],
          "iterations", sa["iterations"], sb["iterations"], flush=True)
    assert sa["rays_tail"] == 0 and sb["rays_tail"] > 0
    # every ray is traced once either way
    assert sb["rays_closest"] + sb["rays_shadow"] + sb["rays_tail"] == sa["rays_closest"] + sa["rays_shadow"]
    assert np.array_equal(wa, wb) and np.array_equal(a, b)
    sc.close()


def test_tail_textured_and_batched(ks, device, cornell):
    """The TEX and BATCH instantiations, the latter through eager batched launches."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).parent))
    from texture_scenes import textured_box

    box = textured_box(stochastic_opacity=False)
    sc = ks.Scene(device, box)
    p = ks.small_pt_params(bounces=6, render_width=96, render_height=96, spp=16, spp_prog_interval=16)
    (a, wa), _ = _render(ks, device, sc, box.camera, p, 96, 96, 0)
    (b, wb), sb = _render(ks, device, sc, box.camera, p, 96, 96, 1 << 30)
    sc.close()
    # the box's cut-out quad makes it an alpha-tested scene: the tail does not apply there
    assert sb["rays_tail"] == 0 and np.array_equal(a, b) and np.array_equal(wa, wb)

    from keepsake_amd.workloads import batch_views

    sc = ks.Scene(device, cornell)
    views = batch_views(ks, cornell.camera, 6, 3)
    out = []
    for tail in (0, 1 << 30):
        device.set_tail(tail)
        films = [ks.Film(device, v.render_width, v.render_height) for _, v in views]
        st = ks.small_pt_render_batch(sc, [(c, v, f) for (c, v), f in zip(views, films)], queue_size=1 << 16,
                                      trace_blocks_per_cu=4, graph=False)
        out.append(([f.download() for f in films], st))
        for f in films:
            f.close()
    device.set_tail(0)
    print("batched tail rays", out[1][1]["rays_tail"], flush=True)
    assert out[0][1]["rays_tail"] == 0 and out[1][1]["rays_tail"] > 0
    for (x, wx), (y, wy) in zip(out[0][0], out[1][0]):
        assert np.array_equal(wx, wy) and np.array_equal(x, y)
    sc.close()


def test_tail_limit_is_validated(ks, device):
    with pytest.raises(ks.KsError):
        device.set_tail(-1)
