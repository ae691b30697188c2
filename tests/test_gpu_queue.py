"""A plain render does not depend on the wavefront's queue size (small_pt_params.queue_size): every work
item's path uses its own Sobol samples and the film folds the per-sample results in sample order, so the
films of one render with a queue far smaller than its paths (generation running through many
iterations, launches bounded by the active paths plus the work left), a queue of no particular power of
two, and the automatic queue (all paths at once) are equal bit for bit, and so are the ray counts."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,seed,bounces", [("cornell", 0, 8), ("spheres_small", 1, 12)])
def test_films_independent_of_queue_size(ks, device, name, seed, bounces):
    hs = ks.procedural_scene(name, seed)
    sc = ks.Scene(device, hs)
    W, H, spp = 160, 96, 16
    cam = dict(hs.camera)
    cam["aspect"] = W / H
    out = []
    for q in (1 << 12, 3 << 14, 0):
        p = ks.small_pt_params(bounces=bounces, render_width=W, render_height=H, spp=spp, spp_prog_interval=spp,
                               queue_size=q)
        f = ks.Film(device, W, H)
        st = ks.small_pt_render(sc, cam, p, f)
        out.append((f.download(), st))
        f.close()
        print(name, "queue", st["queue_size"], "iterations", st["iterations"], "rays", st["rays_closest"],
              st["rays_shadow"], flush=True)
    sc.close()
    (a, wa), sa = out[-1]
    assert sa["queue_size"] == W * H * spp
    for (b, wb), sb in out[:-1]:
        assert sb["rays_closest"] == sa["rays_closest"] and sb["rays_shadow"] == sa["rays_shadow"]
        assert np.array_equal(wa, wb) and np.array_equal(a, b)
