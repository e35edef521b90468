"""Benchmark of the deferred-lighting hot path on one MI355X.

Workload (one step = one frame): 3840x2160, fp16 G-buffer + fp32 depth,
16384 point lights in a 200x30x200 box (range U[1,8]), 32x18x48 log-depth
clusters without a cap.  A frame is the clustered light cull
(hl_light_cull_clusters) followed by deferred shading (hl_deferred_forward),
with every input already resident in HBM.

Prints one JSON line.  value = pixel-light evals per second of whole frames
(Q / frame time; Q = (pixel, light) pairs past the range test, counted on the
GPU once before timing).  The roofline object is for the dominant kernel
(shading), timed with HIP events on the stream it runs on; the CPU baseline
is the C restatement of the reference shaders (oracle/) on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_PIXEL, FLOP_PER_PAIR = 60, 130


def shade_bytes(scene, ncell: int, nlist: int, texel_sum: int) -> int:
    """Algorithmic HBM bytes of one shading launch:
    4 per sky pixel (depth) + (G-buffer hot subset + 16 output) per shaded
    pixel + 4 per cell count and list entry + 56 per light."""
    sky = int((scene.depth >= 1.0).sum())
    shaded = scene.width * scene.height - sky
    return 4 * sky + (texel_sum + 4 + 16) * shaded + 4 * (ncell + nlist) + 56 * len(scene.lights)


def cpu_baseline(scene, params, pool_cap, budget_s: float = 12.0):
    """The oracle (C restatement of light_cull + fs_main, OpenMP) on the host:
    whole frames of the same workload, median time."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    times, q = [], 0
    t_end = time.perf_counter() + budget_s
    reps = 0
    out = np.zeros((scene.height, scene.width, 4), dtype=np.float32)
    while reps < 2 or (time.perf_counter() < t_end and reps < 20):
        t0 = time.perf_counter()
        off, cnt, pool, _ = O.cull_clusters(scene, params, pool_cap)
        grid = O.cluster_grid_host(params, off, cnt, pool)
        _, q = O.shade(scene, grid, out=out)
        times.append(time.perf_counter() - t0)
        reps += 1
    t = statistics.median(times)
    return {"value": q / t, "unit": "pixel-light evals/s", "cores": threads, "kind": "port",
            "sample": f"{reps} whole frames (cluster cull + shade), median {t * 1e3:.1f} ms/frame",
            "ms_per_frame": t * 1e3}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="c3_4k_16k")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; the path shards by screen tiles, see DESIGN.md)",
              file=sys.stderr)
        return 2

    from helio_amd import abi, synth
    from helio_amd.device import Context, DeferredLightPass, LightCullPass, upload_scene

    sc = synth.make_scene(args.workload)
    with Context(0) as ctx:
        L = ctx.lib
        params = synth.cluster_params(L, sc)
        pool_cap = 1 << 25
        ds = upload_scene(ctx, sc)
        cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=pool_cap)
        shade = DeferredLightPass(ctx, sc.width, sc.height)
        cull.prepare(ds)
        shade.prepare(ds)
        cull.execute(ds)
        grid = cull.publish()
        ctx.synchronize()
        if cull.overflow():
            raise RuntimeError("cluster pool overflow")
        q = shade.count_pairs(ds, grid)
        cnt = cull.cgrid.counts.download(np.uint32)
        ncell, nlist = len(cnt), int(cnt.sum())

        for _ in range(args.warmup):
            cull.execute(ds)
            shade.execute(ds, grid)
        ctx.synchronize()
        ev = [(ctx.event(), ctx.event(), ctx.event()) for _ in range(args.steps)]
        t0, t1 = ctx.event(), ctx.event()
        ctx.synchronize()
        ctx.record(t0)
        for a, b, c in ev:
            ctx.record(a)
            cull.execute(ds)
            ctx.record(b)
            shade.execute(ds, grid)
            ctx.record(c)
        ctx.record(t1)
        ctx.synchronize()
        total_ms = ctx.elapsed_ms(t0, t1)
        cull_ms = [ctx.elapsed_ms(a, b) for a, b, _ in ev]
        shade_ms = [ctx.elapsed_ms(b, c) for _, b, c in ev]

    ms_frame = total_ms / args.steps
    shade_avg = sum(shade_ms) / len(shade_ms)
    texel_sum = sc.texel[0] * 2 + sc.texel[1] * 2
    bytes_launch = shade_bytes(sc, ncell, nlist, texel_sum)
    shaded = int((sc.depth < 1.0).sum())
    flops = FLOP_PER_PIXEL * shaded + FLOP_PER_PAIR * q
    gbs = bytes_launch / (shade_avg * 1e-3) / 1e9
    tflops = flops / (shade_avg * 1e-3) / 1e12
    traffic = None
    tfile = ROOT / "profiles" / "traffic_shade_c3_4k_16k.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    res = {
        "metric": "ms/frame & pixel-light evals/s at 4K/16k lights; fwd+adjoint ms at 512²",
        "value": q / (ms_frame * 1e-3),
        "unit": "pixel-light evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_frame,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {sc.width}x{sc.height} fp16 G-buffer + f32 depth, "
                               f"{len(sc.lights)} point lights, clusters "
                               f"{params.grid_x}x{params.grid_y}x{params.grid_z} log-depth, uncapped",
                   "pixels": sc.width * sc.height, "shaded_pixels": shaded, "lights": len(sc.lights),
                   "pairs_q": q, "cluster_entries": nlist, "parallelism": "1 GPU"},
        "ms_per_frame": ms_frame,
        "cull_ms": sum(cull_ms) / len(cull_ms),
        "shade_ms": shade_avg,
        "shade_evals_per_s": q / (shade_avg * 1e-3),
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_shade (forward, clusters, fp16 G-buffer)",
                     "bytes_per_launch": bytes_launch},
        "roofline_valu": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS,
                          "flops_per_launch": flops},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(sc, params, pool_cap)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
