#!/usr/bin/env python3
"""Benchmark: keepsake small_pt on one MI355X.

Workload (one step = one full render): the 64-icosphere scene (1,310,728 triangles, seed 1), 1920x1080,
64 spp, 12 bounces, plain BSDF + NEE sampling, FP32. Prints one JSON line with Mpaths/s, the roofline
of the dominant kernel (closest-hit traversal) measured with HIP events on the render stream, and the
CPU oracle timed on a bounded crop of the same workload on this host's cores.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(1, str(ROOT / "oracle"))  # the CPU oracle, used only by the cpu_baseline leg

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one closest-hit query: ray in (2 x 16 B), hit out (16 B), each node fetch
# (64 B: both child boxes + child pointers), each triangle test (48 B: three vertices + id).
RAY_IO_BYTES = 48
NODE_BYTES = 64
TRI_BYTES = 48

METRIC = "time-to-relMSE 0.01 guided vs plain, Mpaths/s, %HBM peak on 1 MI355X"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(hs, cam, kw, rows: int, threads: int) -> dict:
    import ks_oracle as ko  # CPU oracle: the reported baseline, never the measured GPU path

    t0 = time.time()
    osc = ko.OracleScene.from_host_scene(hs)
    build_s = time.time() - t0
    W, H = kw["render_width"], kw["render_height"]
    y0 = H // 2 - rows // 2
    p = ko.params(**kw, crop_start_x=0, crop_start_y=y0, crop_width=W, crop_height=rows)
    t0 = time.time()
    osc.render(cam, p, threads=threads)
    dt = time.time() - t0
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
    ap.add_argument("--cpu-rows", type=int, default=96)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the path shards by pixel "
            "tiles / sample ranges across GPUs by design, but this build runs on a single MI355X)",
            file=sys.stderr,
        )
        return 2

    from keepsake_amd import api as ks

    seed = 1 if args.scene.startswith("spheres") else 0
    t0 = time.time()
    hs = ks.procedural_scene(args.scene, seed)
    gen_s = time.time() - t0
    log(f"scene {args.scene}: {hs.n_tris} triangles (generated in {gen_s:.1f} s)")
    dev = ks.Device(0)
    t0 = time.time()
    sc = ks.Scene(dev, hs)
    info = sc.info()
    log(f"SBVH: {info['n_nodes']} nodes, {info['n_refs']} refs, depth {info['max_depth']}, "
        f"build {info['build_ms'] / 1e3:.1f} s")
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

    launches = totals["launches"]["closest"]
    k_ms = totals["kernel_ms"]["closest"]
    alg_bytes = (
        RAY_IO_BYTES * totals["rays_closest"]
        + NODE_BYTES * totals["node_visits_closest"]
        + TRI_BYTES * totals["tri_tests_closest"]
    )
    achieved = alg_bytes / (k_ms / 1e3) / 1e9 if k_ms > 0 else 0.0
    traffic = None
    tf = ROOT / "profiles" / "traffic_closest_spheres.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": "k_closest (closest-hit SBVH traversal)",
        "bytes_per_launch": alg_bytes / max(launches, 1),
        "avg_launch_ms": k_ms / max(launches, 1),
    }

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(hs, cam, kw, args.cpu_rows, threads)

    n_rays = (totals["rays_closest"] + totals["rays_shadow"]) / args.steps
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
        },
        "mrays_per_s": round(n_rays / (elapsed / args.steps) / 1e6, 1),
        "kernel_ms_per_step": {k: round(v / args.steps, 2) for k, v in totals["kernel_ms"].items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    film.close()
    sc.close()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
