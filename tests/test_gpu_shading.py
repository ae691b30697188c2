"""Material-sorted shading (ks_device_set_shading(1)): each bounce's paths are split by material class and
shaded by kernels compiled for their classes' BSDF lobes. The films must equal those of the default one-kernel
shading of every path: the classes only drop lobes whose weight is zero for their materials."""

import numpy as np
import pytest

from conftest import rel_mse

pytestmark = pytest.mark.gpu

SCENES = [("cornell", 0, 8, dict()), ("spheres_small", 1, 12, dict()),
          ("room", 3, 16, dict(mis_mode=1, clamp_indirect=0.0))]


@pytest.mark.parametrize("name,seed,bounces,extra", SCENES)
def test_sorted_shading_equals_one_kernel(ks, device, name, seed, bounces, extra):
    hs = ks.procedural_scene(name, seed)
    sc = ks.Scene(device, hs)
    W, H = 160, 96
    cam = dict(hs.camera)
    cam["aspect"] = W / H
    p = ks.small_pt_params(bounces=bounces, render_width=W, render_height=H, spp=16, spp_prog_interval=16, **extra)
    films = []
    for sorted_ in (True, False):
        device.set_shading(sorted_)
        f = ks.Film(device, W, H)
        ks.small_pt_render(sc, cam, p, f)
        films.append(f.download())
        f.close()
    device.set_shading(False)  # the default
    (a, wa), (b, wb) = films
    same = np.all(a == b, axis=-1).mean()
    e = rel_mse(a / np.maximum(wa[..., :1], 1), b / np.maximum(wb[..., :1], 1))
    print(name, "pixels bit-equal", same, "relMSE sorted vs one kernel", e, flush=True)
    assert np.array_equal(wa, wb)
    # the class kernels round a few GGX terms differently (the compiler fuses multiply-adds per kernel):
    # measured 99.6% of pixels bit-equal on the glossy spheres, relMSE 2e-17
    assert same > 0.99 and e < 1e-12
    sc.close()


def test_sorted_shading_guided_and_batched(ks, device, cornell):
    """The guided and the batched wavefronts shade through the sorted kernels too."""
    sc = ks.Scene(device, cornell)
    from keepsake_amd.workloads import batch_views

    views = batch_views(ks, cornell.camera, 6, 3)
    out = []
    for sorted_ in (True, False):
        device.set_shading(sorted_)
        films = [ks.Film(device, p.render_width, p.render_height) for _, p in views]
        ks.small_pt_render_batch(sc, [(c, p, f) for (c, p), f in zip(views, films)], queue_size=1 << 16,
                                 trace_blocks_per_cu=4)
        out.append([f.download() for f in films])
        for f in films:
            f.close()
    device.set_shading(False)  # the default
    for (a, wa), (b, wb) in zip(*out):
        assert np.array_equal(wa, wb)
        assert np.all(a == b, axis=-1).mean() > 0.99
    info = sc.info()
    res = []
    for sorted_ in (True, False):
        device.set_shading(sorted_)
        g = ks.Guide(device, info["bound_lo"], info["bound_hi"], ks.guide_config(seed=1, guide_fraction=0.5))
        f = ks.Film(device, 64, 64)
        ks.small_pt_render_guided(sc, cornell.camera, ks.small_pt_params(bounces=8, render_width=64, render_height=64,
                                  spp=8, spp_prog_interval=8, mis_mode=ks.KS_MIS_CONSISTENT), g, f)
        res.append(f.download())
        f.close()
        g.close()
    device.set_shading(False)  # the default
    # the network's FP32 gradient atomics make two trainings differ in the last bits, so only the (hit,
    # miss) weights, which the camera rays decide, are compared
    assert np.array_equal(res[0][1], res[1][1]) and res[0][1][..., 0].sum() > 0
    sc.close()
