"""The benchmark workloads' inputs that are drawn rather than fixed: the camera views of the batched
render (48 views of the Cornell box at seed-drawn resolutions from {64^2, 128^2, 256x144, 512^2, 1024x576,
1024^2} and spp from {4, 16, 64}, 12 bounces). The benchmark and the tests draw them here, so both see
the same views."""

from __future__ import annotations

import numpy as np

BATCH_RESOLUTIONS = [(64, 64), (128, 128), (256, 144), (512, 512), (1024, 576), (1024, 1024)]
BATCH_SPP = [4, 16, 64]


def batch_views(ks, camera: dict, n_views: int = 48, seed: int = 0, bounces: int = 12):
    """(camera dict, SmallPTParams) per view: resolution and spp drawn from the sets above, the camera
    moved within +-0.2 (x, y) and +-0.3 (z) of the box's front, the view index as the Sobol seed."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_views):
        w, h = BATCH_RESOLUTIONS[rng.integers(len(BATCH_RESOLUTIONS))]
        spp = int(rng.choice(BATCH_SPP))
        cam = dict(camera)
        cam["pos"] = (0.5 + rng.uniform(-0.2, 0.2), 0.5 + rng.uniform(-0.2, 0.2), 2.4 + rng.uniform(-0.3, 0.3))
        cam["aspect"] = w / h
        p = ks.small_pt_params(bounces=bounces, render_width=w, render_height=h, spp=spp, spp_prog_interval=spp,
                               rng_seed=k)
        out.append((cam, p))
    return out
