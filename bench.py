#!/usr/bin/env python3
"""Benchmark: keepsake small_pt on one MI355X.

Workload (one step = one full render): the 64-icosphere scene (1,310,728 triangles, seed 1), 1920x1080,
64 spp, 12 bounces, plain BSDF + NEE sampling, FP32. Prints one JSON line with Mpaths/s, the roofline
of the dominant kernel (closest-hit traversal) measured with HIP events on the render stream, the CPU
oracle timed on a bounded band of the same workload on this host's cores, and the wall time the plain
and the guided tracer need to reach relMSE 0.01 on the slit-lit room against an FP64 reference.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(1, str(ROOT / "oracle"))  # the CPU oracle, used only by the cpu_baseline leg

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one traversal query, as counted by the kernel: ray in (2 x 16 B) and result out (16 B:
# the hit of a closest-hit ray; for an unoccluded shadow ray, 4 B target + 12 B contribution read, the added
# radiance written back is not counted), each node fetch (BVH4: 96 B of child boxes + 16 B of child pointers;
# BVH2: 64 B), each triangle test (48 B: three vertices + id).
RAY_IO_BYTES = 48
NODE_BYTES = {4: 112, 2: 64}
TRI_BYTES = 48
# The implementation-independent form: per closest-hit ray, ray 32 (o, tmin, d, tmax) + hit 20 (t,
# barycentrics, primitive) + 32 per node fetch of the reference SBVH (its node is 32 B) + 52 per triangle test
# (4 B reference + 12 B indices + 36 B vertices), with node fetches and triangle tests counted in the
# reference traversal order; per shadow ray, the same with a 4 B occlusion result in place of the hit.
REF_RAY_BYTES, REF_HIT_BYTES, REF_NODE_BYTES, REF_TRI_BYTES = 32, 20, 32, 52
# Shading, per path vertex (one per closest-hit ray): read hit 20 + path state 64 + material 64 +
# triangle 48 + light record 48 + light CDF 4 per binary-search step; write path state 64 + extension
# ray 32. Per shadow ray: write the ray 32 + its contribution 12. A path's first vertex reads 32 B less:
# its throughput (1, 1, 1) and radiance 0 are constants, not stored state.
SHADE_VERTEX_BYTES = 20 + 64 + 64 + 48 + 48 + 64 + 32
SHADE_SHADOW_BYTES = 32 + 12
SHADE_FIRST_VERTEX_SAVED = 32

METRIC = "time-to-relMSE 0.01 guided vs plain, Mpaths/s, %HBM peak on 1 MI355X"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(hs, cam, kw, rows: int, threads: int, budget_s: float) -> dict:
    """The CPU oracle on a band of rows of the same frame; rows = 0 sizes the band from a short
    calibration run so that the timed render takes about budget_s."""
    import ks_oracle as ko  # CPU oracle: the reported baseline, never the measured GPU path

    t0 = time.time()
    osc = ko.OracleScene.from_host_scene(hs)
    build_s = time.time() - t0
    W, H = kw["render_width"], kw["render_height"]

    def band(n):
        y0 = H // 2 - n // 2
        p = ko.params(**kw, crop_start_x=0, crop_start_y=y0, crop_width=W, crop_height=n)
        t = time.time()
        osc.render(cam, p, threads=threads)
        return y0, time.time() - t

    if rows <= 0:
        _, t_cal = band(16)
        rows = int(min(H, max(16, 16 * budget_s / max(t_cal, 1e-3))))
    y0, dt = band(rows)
    paths = W * rows * kw["spp"]
    osc.close()
    return {
        "value": paths / dt / 1e6,
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"rows {y0}..{y0 + rows - 1} of the {W}x{H} frame at {kw['spp']} spp ({paths} paths), "
        f"{dt:.1f} s render, oracle BVH build {build_s:.1f} s",
    }


BUILDERS = {"host": 0, "device": 1}
BVH_NAME = {0: "host SBVH", 1: "device-built BVH (PLOC)"}


def device_bvh_leg(ks, dev, hs, cam, params, args, host_info) -> dict:
    """Builds the benchmark scene's BVH on the device and renders over it: build time against the host
    SBVH's, and the render rate over each tree."""
    out = {"host_sbvh_build_ms": round(host_info["bvh_build_ms"], 1), "host_sbvh_nodes": host_info["n_nodes"]}
    paths = args.width * args.height * args.spp
    for name in ("device",):
        builds = []
        for _ in range(3):
            sc = ks.Scene(dev, hs, bvh={"branching": args.branching, "builder": BUILDERS[name]})
            info = sc.info()
            builds.append(info["bvh_build_ms"])
            if len(builds) < 3:
                sc.close()
        film = ks.Film(dev, args.width, args.height)
        ks.small_pt_render(sc, cam, params, film)  # warm-up
        times = []
        for _ in range(2):
            film.clear()
            t0 = time.perf_counter()
            ks.small_pt_render(sc, cam, params, film)
            times.append(time.perf_counter() - t0)
        film.close()
        sc.close()
        out[name] = {
            "build_ms": round(min(builds), 2),
            "nodes": info["n_nodes"],
            "depth": info["max_depth"],
            "mpaths_per_s": round(paths / min(times) / 1e6, 2),
        }
    log(f"device BVH: {out}")
    return out


def time_to_relmse(ks, dev, args, scene=("room", 3), size=None, ref_spp=None, max_spp=None, pass_spp=None,
                   methods=(("guided", 0),), builder=0) -> dict:
    """Wall time for the plain and the guided tracer to bring a scene (16 bounces, consistent MIS, no
    clamping) to relMSE <= target against an FP64 plain reference of ref_spp. methods: (name, fp8) of the
    guided runs."""
    hs = ks.procedural_scene(*scene)
    t0 = time.perf_counter()
    sc = ks.Scene(dev, hs, bvh={"builder": builder})
    build_s = time.perf_counter() - t0
    info = sc.info()
    cam = dict(hs.camera)
    W, H = size or (args.room_width, args.room_height)
    cam["aspect"] = W / H
    ref_spp = ref_spp or args.ref_spp
    max_spp = max_spp or args.room_max_spp
    pass_spp = pass_spp or args.plain_pass_spp

    def params(spp, interval):
        return ks.small_pt_params(bounces=16, render_width=W, render_height=H, spp=spp, spp_prog_interval=interval,
                                  mis_mode=ks.KS_MIS_CONSISTENT, clamp_direct=0.0, clamp_indirect=0.0)

    ref = ks.Film(dev, W, H)
    t0 = time.perf_counter()
    st = ks.small_pt_render(sc, cam, params(ref_spp, 256), ref)
    ref_s = time.perf_counter() - t0
    log(f"{scene[0]} reference: {ref_spp} spp in {ref_s:.1f} s ({st['paths'] / ref_s / 1e6:.0f} Mpaths/s)")
    target = args.relmse_target
    res = {"scene": f"{scene[0]} ({hs.n_tris} tris), {W}x{H}, 16 bounces, consistent MIS, no clamping",
           "target": target, "reference_spp": ref_spp, "reference_s": round(ref_s, 1),
           "reference_mpaths_per_s": round(st["paths"] / ref_s / 1e6, 1), "scene_build_s": round(build_s, 2)}

    # one progressive render per method; the progress callback measures relMSE every pass_spp samples
    film = ks.Film(dev, W, H)
    passes = []

    def on_progress(f, spp):
        passes.append((spp, time.perf_counter() - t_start, f.relmse(ref)))

    t_start = time.perf_counter()
    ks.small_pt_render(sc, cam, params(max_spp, pass_spp), film, progress=on_progress)
    hit = next(((s, t, r) for s, t, r in passes if r <= target), None)
    s_last, _, r_last = passes[-1]
    res["plain"] = {"pass_spp": pass_spp, "seconds": round(hit[1], 3) if hit else None,
                    "spp": hit[0] if hit else None, "relmse_at_max_spp": r_last, "max_spp": s_last}
    # the reference's own noise: relMSE(s) = v / s + c over the plain passes (least squares in 1/s); the
    # intercept c is what the ref_spp reference adds to every figure, expected near v / ref_spp
    inv = np.array([1.0 / s for s, _, _ in passes])
    rr = np.array([r for _, _, r in passes])
    v, c = np.linalg.lstsq(np.stack([inv, np.ones_like(inv)], 1), rr, rcond=None)[0]
    res["reference_noise"] = {"fit": "relMSE(s) = v / s + c over the plain passes", "v": float(v), "c": float(c),
                              "expected_c": float(v / ref_spp), "c_over_target": float(c / target)}
    log(f"plain: relMSE {target} at {hit}")

    for name, fp8 in methods:
        cfg = ks.guide_config(seed=1, guide_fraction=args.guide_fraction, train_spp=args.train_spp,
                              max_steps_per_pass=args.guide_steps, max_guided_bounce=args.guide_max_bounce,
                              inference_fp8=fp8)
        guide = ks.Guide(dev, info["bound_lo"], info["bound_hi"], cfg)
        film.clear()
        passes.clear()
        t_start = time.perf_counter()
        ks.small_pt_render_guided(sc, cam, params(max_spp, pass_spp), guide, film, progress=on_progress)
        hit_g = next(((s, t, r) for s, t, r in passes if r <= target), None)
        gst = guide.stats()
        res[name] = {"guide_fraction": args.guide_fraction, "train_spp": args.train_spp,
                     "adam_steps_per_pass": args.guide_steps, "max_guided_bounce": args.guide_max_bounce,
                     "inference": "fp8" if fp8 else "bf16",
                     "seconds": round(hit_g[1], 3) if hit_g else None, "spp": hit_g[0] if hit_g else None,
                     "relmse_at_max_spp": passes[-1][2], "max_spp": passes[-1][0], "adam_steps": gst["steps"],
                     "training_samples": gst["samples"]}
        log(f"{name}: relMSE {target} at {hit_g}")
        if hit and hit_g:
            res[name]["speedup_vs_plain"] = round(hit[1] / hit_g[1], 3)
        guide.close()
    if "guided" in res and "speedup_vs_plain" in res["guided"]:
        res["speedup_guided_vs_plain"] = res["guided"]["speedup_vs_plain"]
    film.close()
    ref.close()
    sc.close()
    return res


def guide_ceiling(ks, dev, args) -> dict:
    """What a trained guide can gain on the room at equal spp: a guide trained for --ceiling-train-spp
    one-spp passes (Adam steps of up to 2^18 samples, at most 8 per pass), then frozen, renders 256 spp
    at guide fraction 0.5, guiding the first 1 / 2 / all bounces; plain renders the same spp. relMSE is
    measured against a 4096-spp FP64 plain reference, whose own noise (v / 4096 from the plain run) is
    subtracted. The ceiling speed-up at equal error is the variance ratio times the ratio of time per
    sample."""
    hs = ks.procedural_scene("room", 3)
    sc = ks.Scene(dev, hs)
    info = sc.info()
    W, H = args.room_width, args.room_height
    cam = dict(hs.camera)
    cam["aspect"] = W / H

    def prm(spp, seed=0):
        return ks.small_pt_params(bounces=16, render_width=W, render_height=H, spp=spp, spp_prog_interval=spp,
                                  mis_mode=ks.KS_MIS_CONSISTENT, clamp_direct=0.0, clamp_indirect=0.0, rng_seed=seed)

    ref = ks.Film(dev, W, H)
    ks.small_pt_render(sc, cam, prm(4096), ref)
    n = 256
    film = ks.Film(dev, W, H)
    t0 = time.perf_counter()
    ks.small_pt_render(sc, cam, prm(n, 1), film)
    t_plain = time.perf_counter() - t0
    r_plain = film.relmse(ref)
    c = r_plain * n / 4096 / (1 + n / 4096)  # the reference's share: v/4096 with v from the plain run
    out = {"spp": n, "reference_spp": 4096, "reference_noise": c, "plain": {"relmse": r_plain, "seconds": t_plain}}
    for mb in (1, 2, 0):
        g = ks.Guide(dev, info["bound_lo"], info["bound_hi"],
                     ks.guide_config(seed=1, guide_fraction=0.5, train_spp=args.ceiling_train_spp, max_steps_per_pass=8,
                                     max_guided_bounce=mb))
        tr = ks.Film(dev, W, H)
        t0 = time.perf_counter()
        ks.small_pt_render_guided(sc, cam, prm(args.ceiling_train_spp, 2), g, tr)
        t_train = time.perf_counter() - t0
        st = g.stats()
        g.set_mode(0.5, 0)  # frozen
        film.clear()
        t0 = time.perf_counter()
        ks.small_pt_render_guided(sc, cam, prm(n, 1), g, film)
        t_g = time.perf_counter() - t0
        r_g = film.relmse(ref)
        vr = (r_plain - c) / max(r_g - c, 1e-12)
        out[f"guided_max_bounce_{mb or 'all'}"] = {
            "relmse": r_g, "seconds": t_g, "training_seconds": t_train, "adam_steps": st["steps"],
            "training_samples": st["samples"], "variance_ratio": round(vr, 3),
            "time_per_spp_ratio": round(t_g / t_plain, 3), "ceiling_speedup": round(vr * t_plain / t_g, 3)}
        tr.close()
        g.close()
    film.close()
    ref.close()
    sc.close()
    log(f"guide ceiling: {out}")
    return out


def batched_views(ks, dev, args) -> dict:
    """48 views of the Cornell box at seed-drawn resolutions and spp (12 bounces) in one batched
    wavefront, against the same views rendered one at a time."""
    from keepsake_amd.workloads import batch_views

    hs = ks.procedural_scene("cornell", 0)
    sc = ks.Scene(dev, hs)
    views = [(cam, p, ks.Film(dev, p.render_width, p.render_height))
             for cam, p in batch_views(ks, hs.camera, args.batch_views, args.batch_seed)]
    paths = sum(v[1].render_width * v[1].render_height * v[1].spp for v in views)
    st = ks.small_pt_render_batch(sc, views)  # first call autotunes (cached afterwards)
    tune_ms = st["wall_ms"]
    for _, _, f in views:
        f.clear()
    t0 = time.perf_counter()
    st = ks.small_pt_render_batch(sc, views)
    batch_s = time.perf_counter() - t0
    for _, _, f in views:
        f.clear()
    t0 = time.perf_counter()
    st_e = ks.small_pt_render_batch(sc, views, queue_size=st["queue_size"],
                                    trace_blocks_per_cu=st["trace_blocks_per_cu"], graph=False)
    eager_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for cam, p, f in views:
        f.clear()
        ks.small_pt_render(sc, cam, p, f)
    single_s = time.perf_counter() - t0
    for _, _, f in views:
        f.close()
    sc.close()
    return {"views": len(views), "paths": paths, "seconds": round(batch_s, 4),
            "mpaths_per_s": round(paths / batch_s / 1e6, 1), "queue_size": st["queue_size"],
            "trace_blocks_per_cu": st["trace_blocks_per_cu"], "first_call_with_autotune_ms": round(tune_ms, 1),
            "eager_launches_seconds": round(eager_s, 4), "one_render_per_view_seconds": round(single_s, 4),
            "speedup_vs_one_render_per_view": round(single_s / batch_s, 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scene", default="spheres")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--bounces", type=int, default=12)
    ap.add_argument("--queue", type=int, default=0)
    ap.add_argument("--branching", type=int, default=4, choices=(2, 4), help="BVH node width on the device")
    ap.add_argument("--builder", default="host", choices=("host", "device"),
                    help="BVH of the timed renders: the host SBVH or a device-built BVH")
    ap.add_argument("--no-device-bvh", action="store_true",
                    help="skip the leg that builds the scene's BVH on the device and renders over it")
    ap.add_argument("--scale-builder", default="device", choices=("host", "device"))
    ap.add_argument("--sorted-shading", action="store_true",
                    help="shade each material class with its own kernel instead of one kernel over every path")
    ap.add_argument("--count-paths", type=int, default=1 << 16,
                    help="random paths whose rays are counted in the reference traversal order (0: skip)")
    ap.add_argument("--cpu-rows", type=int, default=0, help="rows of the CPU band (0: sized for --cpu-seconds)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-relmse", action="store_true", help="skip the guided-vs-plain time-to-relMSE leg")
    ap.add_argument("--room-width", type=int, default=1280)
    ap.add_argument("--room-height", type=int, default=720)
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--room-max-spp", type=int, default=2048)
    ap.add_argument("--pass-spp", dest="plain_pass_spp", type=int, default=8,
                    help="samples per pixel between relMSE checks (plain and guided)")
    ap.add_argument("--relmse-target", type=float, default=0.01)
    ap.add_argument("--guide-fraction", type=float, default=0.3)
    ap.add_argument("--train-spp", type=int, default=8,
                    help="guided renders train during their first N spp (1-spp passes), then render frozen")
    ap.add_argument("--guide-steps", type=int, default=8, help="Adam steps after each training pass")
    ap.add_argument("--guide-max-bounce", type=int, default=1,
                    help="guide only vertices below this bounce (1: the camera rays' hits; 0: every vertex)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 48-view leg")
    ap.add_argument("--guide-ceiling", action="store_true",
                    help="also measure what a long-trained, frozen guide gains at equal spp on the room")
    ap.add_argument("--ceiling-train-spp", type=int, default=512,
                    help="training passes (one spp each) of the guide-ceiling leg")
    ap.add_argument("--fp8-relmse", action="store_true",
                    help="also time the guided tracer with FP8 inference in the room's time-to-relMSE leg")
    ap.add_argument("--no-scale-leg", action="store_true",
                    help="skip the time-to-relMSE leg at scale: the room with 30 icospheres (10M triangles) at "
                    "--scale-width x --scale-height, plain vs guided with FP8 inference")
    ap.add_argument("--scale-width", type=int, default=3840)
    ap.add_argument("--scale-height", type=int, default=2160)
    ap.add_argument("--scale-ref-spp", type=int, default=1024)
    ap.add_argument("--scale-max-spp", type=int, default=512)
    ap.add_argument("--batch-views", type=int, default=48)
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--build-only", type=int, default=0, metavar="N",
                    help="only build the scene's BVH N times with --builder and print the build times")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (calibrates rocprof byte counters)")
    args = ap.parse_args()

    if args.copy_calibration_mib > 0:
        import torch

        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration_mib, "read_bytes": 4 * n, "write_bytes": 4 * n}))
        return 0

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the path shards by pixel "
            "tiles / sample ranges across GPUs by design, but this build runs on a single MI355X)",
            file=sys.stderr,
        )
        return 2

    from keepsake_amd import api as ks

    if args.build_only > 0:
        # the seeds of the benchmark configurations
        hs = ks.procedural_scene(args.scene, {"spheres": 1, "spheres_small": 1, "room": 3, "room_spheres": 4}.get(args.scene, 0))
        dev = ks.Device(0)
        times = []
        for i in range(args.build_only):
            sc = ks.Scene(dev, hs, bvh={"branching": args.branching, "builder": BUILDERS[args.builder]})
            info = sc.info()
            times.append(info["bvh_build_ms"])
            log(f"build {i}: {info['bvh_build_ms']:.2f} ms, {info['n_nodes']} nodes, depth {info['max_depth']}")
            sc.close()
        print(json.dumps({"scene": args.scene, "tris": hs.n_tris, "builder": args.builder, "nodes": info["n_nodes"],
                          "depth": info["max_depth"], "bvh_build_ms": [round(t, 3) for t in times]}), flush=True)
        dev.close()
        return 0

    seed = 1 if args.scene.startswith("spheres") else 0
    t0 = time.time()
    hs = ks.procedural_scene(args.scene, seed)
    gen_s = time.time() - t0
    log(f"scene {args.scene}: {hs.n_tris} triangles (generated in {gen_s:.1f} s)")
    dev = ks.Device(0)
    dev.set_shading(args.sorted_shading)
    t0 = time.time()
    builder = BUILDERS[args.builder]
    sc = ks.Scene(dev, hs, bvh={"branching": args.branching, "builder": builder})
    info = sc.info()
    log(f"{BVH_NAME[builder]}: {info['n_nodes']} BVH{info['branching']} nodes, {info['n_refs']} refs, depth "
        f"{info['max_depth']}, BVH build {info['bvh_build_ms'] / 1e3:.2f} s, scene {info['build_ms'] / 1e3:.1f} s")
    cam = hs.camera
    kw = dict(
        bounces=args.bounces,
        render_width=args.width,
        render_height=args.height,
        spp=args.spp,
        rng_seed=0,
    )
    params = ks.small_pt_params(**kw, spp_prog_interval=args.spp, queue_size=args.queue)
    film = ks.Film(dev, args.width, args.height)

    for i in range(args.warmup):
        film.clear()
        st = ks.small_pt_render(sc, cam, params, film)
        log(f"warmup {i}: {st['wall_ms']:.1f} ms")

    dev.set_timing(True)
    totals = {}
    t_steps = []
    for i in range(args.steps):
        film.clear()
        t0 = time.perf_counter()
        st = ks.small_pt_render(sc, cam, params, film)
        t_steps.append(time.perf_counter() - t0)
        log(f"step {i}: {t_steps[-1] * 1e3:.1f} ms, iterations {st['iterations']}, rays {st['rays_closest']}"
            f"+{st['rays_shadow']}")
        for k, v in st.items():
            if isinstance(v, dict):
                d = totals.setdefault(k, {})
                for kk, vv in v.items():
                    d[kk] = d.get(kk, 0) + vv
            elif isinstance(v, (int, float)):
                totals[k] = totals.get(k, 0) + v
    dev.set_timing(False)
    elapsed = sum(t_steps)
    paths = args.width * args.height * args.spp
    value = paths * args.steps / elapsed / 1e6
    # the same renders again without the per-launch HIP events that the timed steps record for the
    # roofline: their cost is in `value`, not in this figure
    t_untimed = []
    for i in range(args.steps):
        film.clear()
        t0 = time.perf_counter()
        ks.small_pt_render(sc, cam, params, film)
        t_untimed.append(time.perf_counter() - t0)
    untimed = paths * args.steps / sum(t_untimed) / 1e6

    # Implementation-independent traversal bytes: the reference traversal order (SBVH::intersect /
    # intersectBool over the BVH2 SBVH) counted on the host over every ray that a random sample of the
    # workload's paths issues
    ref_order = None
    if builder == BUILDERS["host"] and args.count_paths > 0:
        rng = np.random.default_rng(7)
        m = args.count_paths
        px = rng.integers(0, args.width, m).astype(np.int32)
        py = rng.integers(0, args.height, m).astype(np.int32)
        ps = rng.integers(0, args.spp, m).astype(np.int32)
        t0 = time.time()
        rays_c, rays_s = sc.capture_rays(cam, params, px, py, ps, max_rays=m * (args.bounces + 1))
        nc, tc = sc.traversal_counts(rays_c)
        ns, ts = sc.traversal_counts(rays_s, any_hit=True)
        ref_order = {"paths": m, "closest_rays": len(rays_c), "shadow_rays": len(rays_s),
                     "nodes_per_closest_ray": nc / len(rays_c), "tris_per_closest_ray": tc / len(rays_c),
                     "nodes_per_shadow_ray": ns / len(rays_s), "tris_per_shadow_ray": ts / len(rays_s),
                     "count_s": round(time.time() - t0, 1)}
        log(f"reference-order traversal counts: {ref_order}")

    # k_trace: one launch per wavefront iteration traces that iteration's closest-hit rays and the previous
    # shading pass's shadow rays; a last launch after the loop traces the final pass's shadow rays alone
    launches = totals["launches"]["closest"] + totals["launches"]["shadow"]
    k_ms = totals["kernel_ms"]["closest"] + totals["kernel_ms"]["shadow"]
    node_b = NODE_BYTES[info["branching"]]
    alg_bytes = (
        RAY_IO_BYTES * (totals["rays_closest"] + totals["rays_shadow"])
        + node_b * (totals["node_visits_closest"] + totals["node_visits_shadow"])
        + TRI_BYTES * (totals["tri_tests_closest"] + totals["tri_tests_shadow"])
    )
    achieved = alg_bytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
    traffic = None
    tf = ROOT / "profiles" / "traffic_trace_spheres.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
    avg_ms = k_ms / max(launches, 1)
    kernel_counted = {"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                      "bytes_per_launch": alg_bytes / max(launches, 1),
                      "bytes_per_ray": f"{RAY_IO_BYTES} + {node_b} per BVH{info['branching']} node visit + {TRI_BYTES} "
                      "per triangle test, counted by the kernel"}
    roofline = {
        "bound": "hbm",
        "achieved": kernel_counted["achieved"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kernel_counted["frac"],
        "traffic": traffic,
        "kernel": f"k_trace (closest-hit rays of an iteration + shadow rays of the previous pass, {BVH_NAME[builder]}, "
        f"BVH{info['branching']} nodes)",
        "bytes_per_launch": alg_bytes / max(launches, 1),
        "avg_launch_ms": avg_ms,
        "launches_per_step": launches / args.steps,
    }
    if ref_order is not None:
        # algorithmic bytes with the counts of the reference traversal order over the same rays (the primary
        # figure): per closest-hit ray 32 + 20 + 32 per node fetch + 52 per triangle test; per shadow ray the
        # same with a 4 B result
        per_c = REF_RAY_BYTES + REF_HIT_BYTES + REF_NODE_BYTES * ref_order["nodes_per_closest_ray"] + \
            REF_TRI_BYTES * ref_order["tris_per_closest_ray"]
        per_s = REF_RAY_BYTES + 4 + REF_NODE_BYTES * ref_order["nodes_per_shadow_ray"] + \
            REF_TRI_BYTES * ref_order["tris_per_shadow_ray"]
        ref_bytes = per_c * totals["rays_closest"] + per_s * totals["rays_shadow"]
        ref_achieved = ref_bytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
        roofline.update({"achieved": round(ref_achieved, 1), "frac": round(ref_achieved / HBM_PEAK_GBS, 4),
                         "bytes_per_launch": ref_bytes / max(launches, 1),
                         "bytes_per_closest_ray": round(per_c, 1), "bytes_per_shadow_ray": round(per_s, 1),
                         "basis": "reference traversal order (cuda/sbvh.cu:512-603) over the BVH2 SBVH, counted on the "
                         f"host for the {ref_order['closest_rays']} closest-hit and {ref_order['shadow_rays']} shadow "
                         f"rays of {ref_order['paths']} random paths"})
    roofline["kernel_counted"] = kernel_counted
    if traffic:
        roofline["counter_hbm"] = {"bytes_per_launch": traffic, "achieved": round(traffic / (avg_ms / 1e3) / 1e9, 1),
                                   "frac": round(traffic / (avg_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}

    device_bvh = None
    if not args.no_device_bvh and builder == BUILDERS["host"]:
        device_bvh = device_bvh_leg(ks, dev, hs, cam, params, args, info)

    shade_launches = totals["launches"]["shade"]
    shade_ms = totals["kernel_ms"]["shade"]
    cdf_steps = max(1, math.ceil(math.log2(info["n_lights"] + 1)))
    shade_bytes = (SHADE_VERTEX_BYTES + 4 * cdf_steps) * totals["rays_closest"] + SHADE_SHADOW_BYTES * totals["rays_shadow"] \
        - SHADE_FIRST_VERTEX_SAVED * totals["paths"]
    shade_achieved = shade_bytes / (shade_ms / 1e3) / 1e9 if shade_ms > 0 else 0.0
    roofline_shade = {
        "bound": "hbm",
        "achieved": round(shade_achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(shade_achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": "k_shade (path vertex: emission, NEE, BSDF sampling, roulette)",
        "bytes_per_launch": shade_bytes / max(shade_launches, 1),
        "avg_launch_ms": shade_ms / max(shade_launches, 1),
    }
    tf = ROOT / "profiles" / "traffic_trace_spheres.json"
    if tf.exists():
        roofline_shade["traffic"] = json.loads(tf.read_text()).get("shade_hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        # the host's CPU share of one GPU on the benchmark machine is 16 threads (the box's limit for worker
        # pools, whatever os.cpu_count() reports); --cpu-threads overrides
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(hs, cam, kw, args.cpu_rows, threads, args.cpu_seconds)
        cpu["host_cpus_reported"] = os.cpu_count()

    batch = None
    if not args.no_batch:
        try:
            batch = batched_views(ks, dev, args)
        except ks.KsError as e:
            batch = {"error": str(e)}

    ttr = None
    if not args.no_relmse:
        try:
            methods = (("guided", 0), ("guided_fp8", 1)) if args.fp8_relmse else (("guided", 0),)
            ttr = time_to_relmse(ks, dev, args, methods=methods)
        except ks.KsError as e:  # reported, not hidden: the plain-throughput line stays valid
            ttr = {"error": str(e)}
    ceiling = guide_ceiling(ks, dev, args) if args.guide_ceiling else None
    scale = None
    if not args.no_scale_leg:
        try:
            scale = time_to_relmse(ks, dev, args, scene=("room_spheres", 4), size=(args.scale_width, args.scale_height),
                                   ref_spp=args.scale_ref_spp, max_spp=args.scale_max_spp,
                                   methods=(("guided_fp8", 1),), builder=BUILDERS[args.scale_builder])
        except ks.KsError as e:
            scale = {"error": str(e)}

    n_rays = (totals["rays_closest"] + totals["rays_shadow"] + totals["rays_tail"]) / args.steps
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {
            "workload": f"small_pt plain BSDF+NEE, {args.scene} scene ({hs.n_tris} tris), "
            f"{args.width}x{args.height}, {args.spp} spp, {args.bounces} bounces",
            "scene": args.scene,
            "resolution": [args.width, args.height],
            "spp": args.spp,
            "bounces": args.bounces,
            "queue_size": st["queue_size"],
            "bvh": f"{BVH_NAME[builder]}, BVH{info['branching']} nodes",
        },
        "mrays_per_s": round(n_rays / (elapsed / args.steps) / 1e6, 1),
        "mpaths_per_s_without_timing_events": round(untimed, 1),
        "kernel_ms_per_step": {k: round(v / args.steps, 2) for k, v in totals["kernel_ms"].items()},
        "traversal": {
            "nodes_per_closest_ray": round(totals["node_visits_closest"] / max(totals["rays_closest"], 1), 2),
            "tris_per_closest_ray": round(totals["tri_tests_closest"] / max(totals["rays_closest"], 1), 2),
            "closest_rays_per_step": totals["rays_closest"] // args.steps,
            "shadow_rays_per_step": totals["rays_shadow"] // args.steps,
            "tail_kernel_rays_per_step": totals["rays_tail"] // args.steps,
        },
        "roofline": roofline,
        "roofline_shade": roofline_shade,
        "reference_order_counts": ref_order,
        "cpu_baseline": cpu,
        "time_to_relmse": ttr,
        "time_to_relmse_at_scale": scale,
        "guide_ceiling": ceiling,
        "batched_views": batch,
        "device_bvh": device_bvh,
    }
    print(json.dumps(out), flush=True)
    film.close()
    sc.close()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
