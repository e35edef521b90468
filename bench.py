"""Benchmark of the deferred-lighting hot path on one MI355X.

Headline workload (one step = one frame): 3840x2160, fp16 G-buffer + fp32
depth, 16384 point lights in a 200x30x200 box (range U[1,8]), 32x18x48
log-depth clusters without a cap.  A frame is the clustered light cull
(hl_light_cull_clusters) followed by deferred shading (hl_deferred_forward),
with every input already resident in HBM.

Prints one JSON line.  value = pixel-light evals per second of whole frames
(Q / frame time; Q = (pixel, light) pairs past the range test, counted on the
GPU once before timing).  The roofline object is for the dominant kernel
(shading), timed with HIP events on the stream it runs on; the CPU baseline
is the C restatement of the reference shaders (oracle/) on the host cores.

Further legs, each reported in its own object of the same line:
  batched_4k      the same 4K frame as a one-view batched plan (persistent
                  shading launch), eager and replayed from its HIP graph
  far_4k          the same lights and grid seen from 420 m out with a 90-degree
                  field of view: cluster lists of 0 to 2000+ lights
  fwd_adjoint_512 512x512 fp32 G-buffer, 256 point lights: cluster cull +
                  fused forward / L2 loss / adjoint per step (and forward only)
  gbuffer_pack_4k the G-buffer producer's format contract (hl_gbuffer_pack) at
                  3840x2160 into the reference formats
  shadowed_4k     the 4K/16k frame with shadow_factor (hl_deferred_forward_
                  shadowed): a directional sun (4 cascades) and 8 point + 8
                  spot shadow casters, 1024^2 atlases (60 layers, dynamic and
                  static), PCF (medium preset) and PCSS (high preset), and
                  medium + radiance-cascade GI (hl_deferred_forward_ex)
  forward_lit_4k  forward-lit fragment shading (hl_forward_lit) of the 4K
                  frame's 4.9 M surfaces as fragments (8x8-pixel blocks in
                  raster order) under the same 16384 lights read through
                  entity indices and transforms, on the reference's 16-px tile
                  lists (cap 64) and on the 32x18x48 cluster lists
  reflection_4k   the pass's second draw (hl_deferred_reflection) over the 4K
                  G-buffer: 16 reflection captures over 17 RGBA16F cubes of
                  128^2 x 8 mips, SSR and planar images, RC rough specular
  multiview       the full-size 50-view set (48 probe faces + 2
                  XR eyes at 2064x2208, the eyes with the L2 adjoint) in one
                  autotuned batched plan, graph and eager, against the
                  per-view path on the same grids
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


def roofline_constants() -> dict[str, float]:
    """The #defines of include/helio_hip/roofline.h (the one place the
    algorithmic-work constants and the peaks are fixed)."""
    import re
    text = (ROOT / "include" / "helio_hip" / "roofline.h").read_text()
    return {m.group(1): float(m.group(2)) for m in re.finditer(r"#define\s+(HL_\w+)\s+([0-9.]+)", text)}


RL = roofline_constants()
HBM_PEAK_GBS = RL["HL_PEAK_HBM_GBS"]
FP32_PEAK_TFLOPS = RL["HL_PEAK_FP32_TFLOPS"]
FP32_SCALAR_PEAK_TFLOPS = RL["HL_PEAK_FP32_SCALAR_TFLOPS"]
FLOP_PER_PIXEL, FLOP_PER_PAIR = int(RL["HL_RL_FLOP_PER_PIXEL"]), int(RL["HL_RL_FLOP_PER_PAIR"])


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def shade_bytes(scene, ncell: int, nlist: int, texel_sum: int) -> int:
    """Algorithmic HBM bytes of one shading launch:
    4 per sky pixel (depth) + (G-buffer hot subset + 16 output) per shaded
    pixel + 4 per cell count and list entry + 56 per light."""
    sky = int((scene.depth >= 1.0).sum())
    shaded = scene.width * scene.height - sky
    r = {k: int(v) for k, v in RL.items() if k.startswith("HL_RL_")}
    return (r["HL_RL_SKY_PIXEL_BYTES"] * sky
            + (texel_sum + r["HL_RL_DEPTH_BYTES"] + r["HL_RL_OUTPUT_BYTES"]) * shaded
            + r["HL_RL_CELL_BYTES"] * ncell + r["HL_RL_LIST_ENTRY_BYTES"] * nlist
            + r["HL_RL_LIGHT_BYTES"] * len(scene.lights))


def time_steps(ctx, steps: int, warmup: int, *parts):
    """Runs the callables of `parts` in order per step; returns (mean ms of
    a whole step, per-part mean ms) from HIP events on the context stream.
    The whole-step figure comes from a loop bracketed by two events only (an
    event between parts makes the queue wait, which would add to the step);
    the per-part split from a second loop with an event after every part."""
    for _ in range(warmup):
        for p in parts:
            p()
    ctx.synchronize()
    t0, t1 = ctx.event(), ctx.event()
    ctx.record(t0)
    for _ in range(steps):
        for p in parts:
            p()
    ctx.record(t1)
    ctx.synchronize()
    total = ctx.elapsed_ms(t0, t1) / steps
    if len(parts) == 1:
        return total, [total]
    ev = [[ctx.event() for _ in range(len(parts) + 1)] for _ in range(steps)]
    for e in ev:
        ctx.record(e[0])
        for i, p in enumerate(parts):
            p()
            ctx.record(e[i + 1])
    ctx.synchronize()
    per = [sum(ctx.elapsed_ms(e[i], e[i + 1]) for e in ev) / steps for i in range(len(parts))]
    return total, per


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


def bench_4k(ctx, args):
    from helio_amd import synth
    from helio_amd.batch import BatchPlan, BatchView
    from helio_amd.device import DeferredLightPass, LightCullPass, upload_scene

    sc = synth.make_scene(args.workload)
    params = synth.cluster_params(ctx.lib, sc)
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
    # Q, and the (wavefront, light) visits of the shading walk (each visit
    # evaluates one light for a wavefront's 64 pixels)
    q, visits = shade.walk_stats(ds, grid)
    cnt = cull.cgrid.counts.download(np.uint32)
    ncell, nlist = len(cnt), int(cnt.sum())
    log(f"4K plain: Q={q} wave-light visits={visits} cluster entries={nlist}")
    # the shading call uses the light records the frame's cull built
    from helio_amd import abi
    total, (cull_ms, shade_ms) = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                                            lambda: shade.execute(ds, grid, flags=abi.HL_SHADE_LIGHTS_FROM_CULL))
    res = {"scene": sc, "params": params, "pool_cap": pool_cap, "q": q, "ncell": ncell, "nlist": nlist, "counts": cnt,
           "ms_frame": total, "cull_ms": cull_ms, "shade_ms": shade_ms, "visits": visits}
    # the same frame into the pass's RGBA16F pre_aa target (HL_SHADE_OUT_RGBA16F)
    shade16 = DeferredLightPass(ctx, sc.width, sc.height, out_format=abi.HL_FMT_RGBA16_FLOAT)
    shade16.prepare(ds)
    total16, (_, shade16_ms) = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                                          lambda: shade16.execute(ds, grid, flags=abi.HL_SHADE_LIGHTS_FROM_CULL))
    res["rgba16f"] = {"ms_per_frame": total16, "shade_ms": shade16_ms, "evals_per_s": q / (total16 * 1e-3)}
    log(f"4K into RGBA16F: {total16:.4f} ms/frame (shade {shade16_ms * 1e3:.1f} us)")
    if not args.no_batched:
        out = ctx.alloc(sc.width * sc.height * 16)
        # the plain path's cluster grid, so both legs build the same lists
        plan = BatchPlan(ctx, ds.lights, len(sc.lights), [BatchView(scene=ds, subset=None,
                                                                     num_lights=len(sc.lights), out=out,
                                                                     clusters=params)])
        try:
            eager, _ = time_steps(ctx, args.steps, args.warmup, plan.launch)
            plan.capture()
            graph, _ = time_steps(ctx, args.steps, args.warmup, plan.launch)
        finally:
            plan.destroy()
        res["batched"] = {"ms_per_frame_graph": graph, "ms_per_frame_eager": eager,
                          "evals_per_s": q / (graph * 1e-3),
                          "note": "one-view plan on the same cluster grid: batched cull + persistent shading launch"}
        log(f"4K batched: graph {graph:.3f} ms, eager {eager:.3f} ms")
    return res


def bench_far(ctx, args):
    """The 4K/16k lights and grid seen from 420 m out with a 90-degree field of
    view (c3_4k_16k_far): cluster lists of 0 to 2000+ lights."""
    import argparse
    a = argparse.Namespace(**vars(args))
    a.workload, a.no_batched = "c3_4k_16k_far", True
    r = bench_4k(ctx, a)
    cnt = r["counts"]
    log(f"4K far camera: {r['ms_frame']:.4f} ms/frame (shade {r['shade_ms'] * 1e3:.1f} us), Q={r['q']}, "
        f"longest list {int(cnt.max())}")
    return {"workload": "c3_4k_16k_far: the 4K/16k lights and 32x18x48 grid from (0, 60, 420) with a 90-degree "
                        "field of view", "ms_per_frame": r["ms_frame"], "cull_ms": r["cull_ms"],
            "shade_ms": r["shade_ms"], "pairs_q": r["q"], "evals_per_s": r["q"] / (r["ms_frame"] * 1e-3),
            "cluster_entries": r["nlist"], "longest_list": int(cnt.max()),
            "clusters_2000_plus": int((cnt >= 2000).sum())}


def bench_gbuffer_pack(ctx, args):
    """hl_gbuffer_pack at 3840x2160 into the reference formats (RGBA8 albedo/orm,
    RGBA16F normal/emissive), metallic workflow: 64 B read + 24 B written per pixel."""
    from helio_amd.device import pack_gbuffer, surface_desc
    from helio_amd import abi
    W, H = 3840, 2160
    rng = np.random.default_rng(5)
    plane = rng.uniform(0.0, 1.0, size=(H, W, 4)).astype(np.float32)
    planes = [ctx.upload(plane) for _ in range(4)]
    outs = [ctx.alloc(W * H * tb) for tb in (4, 8, 4, 8)]
    g = abi.hl_gbuffer()
    g.width, g.height = W, H
    for t, o, fmt, tb in zip((g.albedo, g.normal, g.orm, g.emissive), outs,
                             (abi.HL_FMT_RGBA8_UNORM, abi.HL_FMT_RGBA16_FLOAT, abi.HL_FMT_RGBA8_UNORM,
                              abi.HL_FMT_RGBA16_FLOAT), (4, 8, 4, 8)):
        t.data_d, t.pitch_bytes, t.format = o.ptr, W * tb, fmt
    s = surface_desc(W, H, *planes)
    ms, _ = time_steps(ctx, args.steps, args.warmup, lambda: pack_gbuffer(ctx, s, g))
    nbytes = W * H * (4 * 16 + 4 + 8 + 4 + 8)
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"G-buffer pack 4K: {ms * 1e3:.1f} us, {gbs:.0f} GB/s")
    return {"ms_per_pack": ms, "bytes_per_pack": nbytes, "achieved_GBs": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS,
            "workload": "3840x2160 RGBA32F surface planes -> RGBA8 albedo/orm + RGBA16F normal/emissive "
                        "(metallic-workflow F0), hl_gbuffer_pack"}


def bench_shadowed(ctx, args):
    """Cluster cull + shadowed shading of the 4K/16k frame with 17 shadow
    casters; whole frames timed as in the headline, per preset."""
    from helio_amd import synth
    from helio_amd.device import DeferredLightPass, DeviceShadows, LightCullPass, upload_scene
    sc = synth.make_scene(args.workload)
    sh = synth.add_shadow_casters(sc, n_point=8, n_spot=8, directional=True, size=1024, quality="medium")
    ds = upload_scene(ctx, sc)
    dsh = DeviceShadows(ctx, sh)
    params = synth.cluster_params(ctx.lib, sc)
    cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 24)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    shade.prepare(ds)
    cull.execute(ds)
    grid = cull.publish()
    q = shade.count_pairs(ds, grid)
    res = {"workload": f"{args.workload} with a directional sun (4 cascades) and 8 point + 8 spot shadow casters: "
                       f"{sh.layers} layers of {sh.size}^2 in each atlas (dynamic, static)",
           "pairs_q": q, "layers": sh.layers, "atlas_size": sh.size}
    ms_plain, parts = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                                 lambda: shade.execute(ds, grid))
    res["unshadowed"] = {"ms_per_frame": ms_plain, "shade_ms": parts[1]}
    for quality in ("medium", "high"):
        dsh.desc.config = synth.shadow_config(quality)
        ms, parts = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                               lambda: shade.execute(ds, grid, shadows=dsh))
        res[quality] = {"ms_per_frame": ms, "shade_ms": parts[1], "evals_per_s": q / (ms * 1e-3),
                        "pcss": bool(dsh.desc.config.enable_pcss),
                        "pcf_samples": int(dsh.desc.config.pcf_sample_count)}
        log(f"shadowed 4K {quality}: {ms:.3f} ms/frame (shade {parts[1] * 1e3:.1f} us)")
    # radiance-cascade indirect diffuse on top of the medium shadows (hl_deferred_forward_ex)
    from helio_amd.device import DeviceTexture
    atlas = synth.make_rc_probes(sc, box=((-100.0, -1.0, -100.0), (100.0, 30.0, 100.0)))
    rc = DeviceTexture(ctx, atlas)
    shade.prepare(ds)  # the globals now carry has_rc_gi and the probe box
    dsh.desc.config = synth.shadow_config("medium")
    ms, parts = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                           lambda: shade.execute(ds, grid, shadows=dsh, rc=rc))
    res["medium_rc_gi"] = {"ms_per_frame": ms, "shade_ms": parts[1], "evals_per_s": q / (ms * 1e-3),
                           "rc_atlas": "64x1024 RGBA16F, probe box (-100,-1,-100)..(100,30,100)"}
    log(f"shadowed 4K medium + RC GI: {ms:.3f} ms/frame (shade {parts[1] * 1e3:.1f} us)")
    return res


def bench_reflection(ctx, args):
    """The pass's second draw (hl_deferred_reflection) over the 4K/16k frame:
    16 captures (8 oriented boxes, 8 spheres, one not resident) over a cube
    array of 17 RGBA16F cubes of 128^2 with 8 mips, full-resolution SSR and
    planar images, and the radiance-cascade rough-specular fallback."""
    from helio_amd import abi, synth
    from helio_amd.device import DeviceReflections, ReflectionPass, upload_scene
    sc = synth.make_scene(args.workload)
    refl = synth.make_reflections(sc, n_box=8, n_sphere=8, face=128, mips=8, rc=True)
    ds = upload_scene(ctx, sc)
    dr = DeviceReflections(ctx, refl)
    out = ctx.alloc(sc.width * sc.height * 16)
    out.fill_u32(0)
    rp = ReflectionPass(ctx)
    ms, _ = time_steps(ctx, args.steps, args.warmup, lambda: rp.execute(ds, dr, out))
    shaded = int((sc.depth < 1.0).sum())
    texel = sc.texel[0] + 2 * sc.texel[1]  # orm, normal, emissive
    # per shaded pixel: depth, G-buffer texels, SSR and planar (RGBA16F), output read and written;
    # per sky pixel: depth.  The cube array (17.8 MB) stays in the caches.
    nbytes = shaded * (4 + texel + 8 + 8 + 32) + (sc.width * sc.height - shaded) * 4
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"reflection 4K: {ms * 1e3:.1f} us, {gbs:.0f} GB/s")
    return {"ms_per_pass": ms, "bytes_per_pass": nbytes, "achieved_GBs": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS,
            "cube_array_bytes": int(refl.cubes.nbytes),
            "workload": f"{args.workload} G-buffer; 16 reflection captures (8 box, 8 sphere, 1 without a cube), "
                        "17 RGBA16F cubes of 128^2 x 8 mips, RGBA16F SSR + planar, RC rough specular; "
                        "hl_deferred_reflection added over the frame"}


def bench_forward_lit(ctx, args):
    """hl_forward_lit over every surface of the 4K/16k scene as one fragment
    (80-byte records), lights positioned through entity indices and
    transforms, on the tile lists and on the cluster lists."""
    from helio_amd import synth
    from helio_amd.device import ForwardLitPass, LightCullPass, upload_scene
    sc = synth.make_scene(args.workload)
    sc.lights, entity, transforms = synth.make_light_transforms(sc.lights, seed=4)
    ds = upload_scene(ctx, sc)
    frags = synth.make_fragments(sc, "tiles")
    n = len(frags)
    d_fr, d_ent, d_tr = ctx.upload(frags), ctx.upload(entity), ctx.upload(transforms)
    fwd = ForwardLitPass(ctx, sc.width, sc.height)
    fwd.prepare(ds)
    res = {"workload": f"{n} fragments (every surface of the {sc.width}x{sc.height} frame, 8x8-pixel blocks "
                       f"in raster order, 80 B each) under {len(sc.lights)} lights read through entity "
                       f"indices and transforms; RGBA32F per fragment",
           "fragments": n}
    params = synth.cluster_params(ctx.lib, sc)
    for mode, cull in (("tiles", LightCullPass(ctx, sc.width, sc.height)),
                       ("clusters", LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 24))):
        cull.prepare(ds)
        cull.execute(ds)
        cull_ms, _ = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds))
        grid = cull.publish()
        kw = dict(entity=d_ent, transforms=d_tr, num_transforms=len(transforms))
        pairs = fwd.count_pairs(ds, d_fr, n, grid, **kw)
        ms, _ = time_steps(ctx, args.steps, args.warmup, lambda: fwd.execute(ds, d_fr, n, grid, **kw))
        res[mode] = {"ms_per_call": ms, "pairs": pairs, "evals_per_s": pairs / (ms * 1e-3),
                     "fragments_per_s": n / (ms * 1e-3),
                     "bytes_fragment_io": 96 * n, "fragment_io_GBs": 96 * n / (ms * 1e-3) / 1e9,
                     "cull_ms": cull_ms}
        log(f"forward-lit 4K {mode}: {ms * 1e3:.1f} us, {pairs / (ms * 1e-3) / 1e9:.1f} G evals/s; "
            f"its light cull {cull_ms * 1e3:.1f} us")
    res["tiles"]["grid"] = "reference tile_light_lists (16 px, cap 64, ascending)"
    res["clusters"]["grid"] = f"{params.grid_x}x{params.grid_y}x{params.grid_z} log-depth clusters, uncapped"
    return res


def bench_adjoint(ctx, args):
    from helio_amd import synth
    from helio_amd.device import AdjointBuffers, DeferredLightPass, LightCullPass, upload_scene

    gt = synth.make_scene("c4_512_adj")
    guess = synth.perturb_for_adjoint(gt)
    params = synth.cluster_params(ctx.lib, gt)
    pool_cap = 1 << 22
    # target: the forward render of the ground truth
    ds_gt = upload_scene(ctx, gt)
    cull_gt = LightCullPass(ctx, gt.width, gt.height, clusters=params, pool_cap=pool_cap)
    shade_gt = DeferredLightPass(ctx, gt.width, gt.height)
    cull_gt.prepare(ds_gt)
    shade_gt.prepare(ds_gt)
    cull_gt.execute(ds_gt)
    shade_gt.execute(ds_gt, cull_gt.publish())
    target = shade_gt.out
    ds = upload_scene(ctx, guess)
    cull = LightCullPass(ctx, gt.width, gt.height, clusters=params, pool_cap=pool_cap)
    shade = DeferredLightPass(ctx, gt.width, gt.height)
    cull.prepare(ds)
    shade.prepare(ds)
    grid = cull.publish()
    cull.execute(ds)
    q, visits = shade.walk_stats(ds, grid)
    grads = AdjointBuffers(ctx, gt.width, gt.height, len(guess.lights))
    fb, (fb_cull, fb_shade) = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                                         lambda: shade.forward_backward_l2(ds, grid, target, grads))
    fw, (_, fw_shade) = time_steps(ctx, args.steps, args.warmup, lambda: cull.execute(ds),
                                   lambda: shade.execute(ds, grid))
    loss = float(grads.loss.download(np.float64)[0])
    log(f"512 adjoint: fwd+adjoint {fb:.4f} ms/step, forward {fw:.4f} ms/step, loss {loss:.4g}, Q={q}")
    # algorithmic bytes of the fused launch: per pixel the RGBA32F G-buffer
    # (4 x 16 B) + depth (4), the target read (16), the output written (16)
    # and the five gradient planes (20); per light its record (80 B) and the
    # gradient slots (kGradSlots = 8 copies of 4 doubles)
    npx, nl = gt.width * gt.height, len(guess.lights)
    nbytes = npx * (64 + 4 + 16 + 16 + 20) + nl * (80 + 8 * 32)
    gbs = nbytes / (fb_shade * 1e-3) / 1e9
    return {"ms_per_step": fb, "cull_ms": fb_cull, "fused_forward_adjoint_ms": fb_shade,
            "forward_only_ms_per_step": fw, "forward_shade_ms": fw_shade, "loss": loss,
            "pairs_q": q, "wave_light_visits": visits, "lane_utilisation": q / (64 * visits) if visits else None,
            "fused_evals_per_s": q / (fb_shade * 1e-3),
            "fused_roofline": {"bound": "latency (VALU + FP64 atomics at 3 waves per SIMD)", "achieved": gbs,
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                               "bytes_per_launch": nbytes,
                               "kernel": "k_shade (fused forward + L2 adjoint, clusters, fp32 G-buffer)"},
            "workload": "512x512 fp32 G-buffer, 256 point lights, clusters "
                        f"{params.grid_x}x{params.grid_y}x{params.grid_z}; per step: cluster cull + "
                        "forward + L2 loss + gradients wrt albedo/roughness/metallic per pixel and "
                        "position/intensity per light"}


def bench_multiview(ctx, args):
    """The full-size 50-view set (48 probe faces at 128^2, 256^2 and 512^2, two
    2064x2208 XR eyes with the L2 adjoint; 8192-light pool, per-view light
    subsets) in one autotuned batched plan (one launch chain, one captured
    graph), against the per-view path the reference takes (each view culled
    and shaded with its own calls: crates/helio/src/renderer/render.rs:804-873)."""
    from helio_amd import synth
    from helio_amd.batch import BatchPlan, make_multiview
    from helio_amd.device import AdjointBuffers, DeferredLightPass, LightCullPass

    pool, pool_size, views, hosts, _ = make_multiview(ctx, small=False)
    for v in views:
        if v.adjoint:
            t = np.zeros((v.scene.host.height, v.scene.host.width, 4), np.float32)
            t[..., :3] = 0.25
            v.target.upload(t)
    log("50 views full size: autotuning the plan")
    plan = BatchPlan(ctx, pool, pool_size, views, autotune=True, max_trials=args.tune_trials)
    try:
        cfg = {i: plan.config(i) for i in range(len(views))}
        eager, _ = time_steps(ctx, args.steps, args.warmup, plan.launch)
        plan.capture()
        graph, _ = time_steps(ctx, args.steps, args.warmup, plan.launch)
    finally:
        plan.destroy()
    # the per-view path on the same cluster grids
    calls = []
    for i, v in enumerate(views):
        sc, ds = v.scene.host, v.scene
        c = cfg[i]
        params = synth.cluster_params(ctx.lib, sc, grid=(c.grid_x, c.grid_y, c.grid_z), cap=0)
        cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 24)
        shade = DeferredLightPass(ctx, sc.width, sc.height)
        cull.prepare(ds)
        shade.prepare(ds)
        grid = cull.publish()
        calls.append(lambda cull=cull, ds=ds: cull.execute(ds))
        if v.adjoint:
            grads = AdjointBuffers(ctx, sc.width, sc.height, v.num_lights)
            calls.append(lambda shade=shade, ds=ds, grid=grid, t=v.target, g=grads:
                         shade.forward_backward_l2(ds, grid, t, g))
        else:
            calls.append(lambda shade=shade, ds=ds, grid=grid: shade.execute(ds, grid))
    per_view, _ = time_steps(ctx, args.steps, args.warmup, lambda: [f() for f in calls])
    px = sum(v.scene.host.width * v.scene.host.height for v in views)
    log(f"50 views full size: one plan graph {graph:.3f} ms, eager {eager:.3f} ms; per view {per_view:.3f} ms")
    return {"workload": "48 probe faces (16 each at 128^2, 256^2, 512^2) + 2 eyes at 2064x2208 with the L2 "
                        "adjoint, 8192-light pool, log-uniform subsets of 32..4096 lights, reference G-buffer "
                        "formats", "views": len(views), "pixels": px, "plans": 1,
            "ms_per_frame_graph": graph, "ms_per_frame_eager": eager, "ms_per_frame_per_view": per_view,
            "speedup_vs_per_view": per_view / graph,
            "eye_grids": [f"{cfg[i].grid_x}x{cfg[i].grid_y}x{cfg[i].grid_z}" for i, v in enumerate(views)
                          if v.adjoint]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the one-view batched 4K leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 512x512 forward+adjoint leg")
    ap.add_argument("--no-pack", action="store_true", help="skip the 4K G-buffer pack leg")
    ap.add_argument("--no-forward", action="store_true", help="skip the 4K forward-lit leg")
    ap.add_argument("--no-shadows", action="store_true", help="skip the 4K shadowed leg")
    ap.add_argument("--no-reflection", action="store_true", help="skip the 4K reflection-composite leg")
    ap.add_argument("--no-far", action="store_true", help="skip the 4K far-camera (long cluster lists) leg")
    ap.add_argument("--no-multiview", action="store_true",
                    help="skip the full-size 50-view leg (one autotuned batched plan against the per-view path)")
    ap.add_argument("--tune-trials", type=int, default=4)
    ap.add_argument("--workload", default="c3_4k_16k")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (a kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; the path would shard by screen tiles and views)",
              file=sys.stderr)
        return 2

    from helio_amd.device import Context

    if args.calibrate_copy:
        with Context(0) as ctx:
            n = args.calibrate_copy << 20
            a, b = ctx.alloc(n), ctx.alloc(n)
            a.fill_u32(0x3F800000)
            for _ in range(5):
                ctx.lib.call("hl_memcpy_dtod", ctx.h, b.ptr, a.ptr, n)
            ctx.synchronize()
        print(json.dumps({"calibration_copy_bytes": n, "copies": 5}), flush=True)
        return 0

    with Context(0) as ctx:
        r4k = bench_4k(ctx, args)
        far = None if args.no_far else bench_far(ctx, args)
        adj = None if args.no_adjoint else bench_adjoint(ctx, args)
        pack = None if args.no_pack else bench_gbuffer_pack(ctx, args)
        fwd = None if args.no_forward else bench_forward_lit(ctx, args)
        shd = None if args.no_shadows else bench_shadowed(ctx, args)
        rfl = None if args.no_reflection else bench_reflection(ctx, args)
        mv = None if args.no_multiview else bench_multiview(ctx, args)

    sc, q = r4k["scene"], r4k["q"]
    ms_frame, shade_ms = r4k["ms_frame"], r4k["shade_ms"]
    texel_sum = sc.texel[0] * 2 + sc.texel[1] * 2
    bytes_launch = shade_bytes(sc, r4k["ncell"], r4k["nlist"], texel_sum)
    shaded = int((sc.depth < 1.0).sum())
    flops = FLOP_PER_PIXEL * shaded + FLOP_PER_PAIR * q
    gbs = bytes_launch / (shade_ms * 1e-3) / 1e9
    tflops = flops / (shade_ms * 1e-3) / 1e12
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE,
    # calibrated against a copy of known size (the summary states both runs)
    traffic = None
    tfile = ROOT / "profiles" / "traffic_shade_c3_4k_16k.md"
    if tfile.exists():
        for line in tfile.read_text().splitlines():
            if line.startswith("bytes_per_launch = "):
                traffic = int(line.split("=")[1])
    params = r4k["params"]
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
                   "pairs_q": q, "cluster_entries": r4k["nlist"], "parallelism": "1 GPU"},
        "ms_per_frame": ms_frame,
        "cull_ms": r4k["cull_ms"],
        "shade_ms": shade_ms,
        "shade_evals_per_s": q / (shade_ms * 1e-3),
        "walk": {"wave_light_visits": r4k["visits"], "lane_utilisation": q / (64 * r4k["visits"]),
                 "note": "each visit evaluates one light for a wavefront's 64 pixels; Q / (64 visits) is the "
                         "share of those lanes that hold an in-range pixel"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_shade (forward, clusters, fp16 G-buffer)",
                     "bytes_per_launch": bytes_launch},
        "roofline_valu": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS,
                          "frac_of_scalar_fp32": tflops / FP32_SCALAR_PEAK_TFLOPS,
                          "flops_per_launch": flops},
    }
    res["rgba16f_target_4k"] = r4k["rgba16f"]
    if "batched" in r4k:
        res["batched_4k"] = r4k["batched"]
    if far is not None:
        res["far_4k"] = far
    if adj is not None:
        res["fwd_adjoint_512"] = adj
    if pack is not None:
        res["gbuffer_pack_4k"] = pack
    if fwd is not None:
        res["forward_lit_4k"] = fwd
    if shd is not None:
        res["shadowed_4k"] = shd
    if rfl is not None:
        res["reflection_4k"] = rfl
    if mv is not None:
        res["multiview"] = mv
    else:
        res["multiview"] = {"skipped": "--no-multiview"}
    if not args.no_cpu:
        log("CPU baseline")
        res["cpu_baseline"] = cpu_baseline(sc, params, r4k["pool_cap"])
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
