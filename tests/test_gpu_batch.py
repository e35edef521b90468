"""Batched multi-view plan on the GPU (48 probe faces + 2 XR eyes, each view
shrunk 8x per side; 8192-light pool; log-uniform [32, 4096] subsets; L2
adjoint on the two eyes).

Bars:
  * every view's radiance equals the single-view path run with the same
    cluster grid bit for bit (same kernels, same lists, same order); an
    adjoint view whose wave tile differs from the single-view path's 8 x 8
    within 1e-6 relative (the fused forward adds its four waves' partial
    sums, each over every fourth light the wave tile's walk visits, so the
    tile shape and the cluster grid move the last bits);
  * the eyes' gradients equal the single-view fused adjoint within 1e-5
    relative + 1e-7 of the largest magnitude (FP64 atomics reorder sums);
  * each replay of the captured HIP graph equals the eager frame, with eager
    work interleaved between replays;
  * the autotuner returns a valid configuration and its cache round-trips;
  * views given a fixed cluster grid (with a per-cluster cap) get exactly that
    grid, and equal the single-view path on it as above; the last view is
    in that plan twice with two grids (both write its buffers), so its
    fused output is checked within 1e-6 relative.
"""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.batch import BatchPlan, BatchView, make_multiview
from helio_amd.device import ClusterGrid, LightCullPass

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mv(ctx):
    pool, pool_size, views, hosts, pool_lights = make_multiview(ctx, small=True)
    # targets of the eyes: 0.9 x a forward render
    for v in views:
        if v.adjoint:
            t = np.zeros((v.scene.host.height, v.scene.host.width, 4), np.float32)
            t[..., :3] = 0.25
            v.target.upload(t)
    return pool, pool_size, views, hosts, pool_lights


def single_view(ctx, hlib, v, cfg, cap=0):
    """The single-view path on the plan's cluster grid."""
    sc = v.scene.host
    sub_lights = sc.lights
    p = synth.cluster_params(hlib, sc, grid=(cfg.grid_x, cfg.grid_y, cfg.grid_z), cap=cap)
    lights = ctx.upload(sub_lights) if len(sub_lights) else ctx.alloc(128)
    cull = LightCullPass(ctx, sc.width, sc.height, clusters=p, pool_cap=1 << 22)
    cam = v.scene.camera
    ctx.lib.call("hl_light_cull_clusters", ctx.h, cam.ptr, C.byref(p), lights.ptr, len(sub_lights),
                 cull.cgrid.offsets.ptr, cull.cgrid.counts.ptr, cull.cgrid.pool.ptr, cull.cgrid.pool_cap,
                 cull.cgrid.overflow.ptr)
    grid = cull.publish()
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 16)
    res = {}
    if not v.adjoint:
        ctx.lib.call("hl_deferred_forward", ctx.h, cam.ptr, C.byref(sc.globals), C.byref(v.scene.gbuffer),
                     lights.ptr, len(sub_lights), C.byref(grid), out.ptr, W * 16, 0)
    else:
        n = len(sub_lights)
        loss = ctx.alloc(8)
        dpx = ctx.alloc(5 * W * H * 4)
        dpos = ctx.alloc(3 * max(n, 1) * 4)
        dint = ctx.alloc(max(n, 1) * 4)
        ctx.lib.call("hl_deferred_forward_backward_l2", ctx.h, cam.ptr, C.byref(sc.globals),
                     C.byref(v.scene.gbuffer), lights.ptr, n, C.byref(grid), v.target.ptr, W * 16, out.ptr,
                     W * 16, loss.ptr, dpx.ptr, dpos.ptr, dint.ptr)
        res.update(loss=float(loss.download(np.float64)[0]), dpx=dpx.download(np.float32),
                   dpos=dpos.download(np.float32)[: 3 * n].reshape(n, 3), dint=dint.download(np.float32)[:n])
    res["out"] = out.download(np.float32, (H, W, 4))
    return res


def read_view(v, pool_size, subset):
    sc = v.scene.host
    r = {"out": v.out.download(np.float32, (sc.height, sc.width, 4))}
    if v.adjoint:
        r["loss"] = float(v.loss.download(np.float64)[0])
        r["dpx"] = v.d_pixel.download(np.float32)
        r["dpos"] = v.d_pos.download(np.float32).reshape(pool_size, 3)[subset]
        r["dint"] = v.d_int.download(np.float32)[subset]
    return r


def close(a, b, rel=1e-5):
    scale = np.abs(b).max() if b.size else 0.0
    return np.all(np.abs(a - b) <= rel * np.abs(b) + 1e-7 * scale)


def test_batch_equals_single_view(ctx, hlib, mv):
    pool, pool_size, views, hosts, _ = mv
    print("[batch] 50 views eager vs single-view", flush=True)
    plan = BatchPlan(ctx, pool, pool_size, views)
    try:
        for v in views:
            v.out.fill_u32(0)
        plan.launch()
        ctx.synchronize()
        got = [read_view(v, pool_size, sub) for v, (_, sub) in zip(views, hosts)]
        for i, v in enumerate(views):
            cfg = plan.config(i)
            ref = single_view(ctx, hlib, v, cfg)
            sky = v.scene.host.depth >= 1.0
            if v.adjoint and (cfg.wave_tile_w, cfg.wave_tile_h) != (8, 8):
                assert close(got[i]["out"][~sky], ref["out"][~sky], rel=1e-6), f"view {i}"
            else:
                assert np.array_equal(got[i]["out"][~sky], ref["out"][~sky]), f"view {i}"
            if v.adjoint:
                assert abs(got[i]["loss"] - ref["loss"]) <= 1e-6 * abs(ref["loss"])
                assert close(got[i]["dpx"], ref["dpx"]), f"view {i} pixel grads"
                assert close(got[i]["dpos"], ref["dpos"]), f"view {i} light pos grads"
                assert close(got[i]["dint"], ref["dint"]), f"view {i} intensity grads"
    finally:
        plan.destroy()


def test_one_view_plans_equal_single_view(ctx, hlib, mv):
    """Plans of one view (the shading launch takes the view's arguments as a
    kernel argument): a 512^2 face (forward) bit for bit, an eye (fused
    forward + L2 adjoint) as in test_batch_equals_single_view; eager and
    replayed from the captured graph."""
    pool, pool_size, views, hosts, _ = mv
    print("[batch] one-view plans vs single-view", flush=True)
    for i in (40, len(views) - 1):
        v, (_, sub) = views[i], hosts[i]
        plan = BatchPlan(ctx, pool, pool_size, [v])
        try:
            cfg = plan.config(0)
            ref = single_view(ctx, hlib, v, cfg)
            sky = v.scene.host.depth >= 1.0
            for rep in range(3):
                v.out.fill_u32(0)
                if rep == 1:
                    plan.capture()
                plan.launch()
                ctx.synchronize()
                got = read_view(v, pool_size, sub)
                if v.adjoint:
                    assert close(got["out"][~sky], ref["out"][~sky], rel=1e-6), f"view {i} rep {rep}"
                    assert abs(got["loss"] - ref["loss"]) <= 1e-6 * abs(ref["loss"]), f"view {i} rep {rep}"
                    assert close(got["dpx"], ref["dpx"]), f"view {i} rep {rep} pixel grads"
                    assert close(got["dpos"], ref["dpos"]), f"view {i} rep {rep} light pos grads"
                    assert close(got["dint"], ref["dint"]), f"view {i} rep {rep} intensity grads"
                else:
                    assert np.array_equal(got["out"][~sky], ref["out"][~sky]), f"view {i} rep {rep}"
        finally:
            plan.destroy()


def test_graph_replays_match_eager(ctx, mv):
    pool, pool_size, views, hosts, _ = mv
    print("[batch] graph replays", flush=True)
    plan = BatchPlan(ctx, pool, pool_size, views)
    try:
        plan.launch()
        ctx.synchronize()
        eager = [read_view(v, pool_size, sub) for v, (_, sub) in zip(views, hosts)]
        plan.capture()
        for rep in range(3):
            for v in views:  # eager work between replays
                v.out.fill_u32(0x7fc00000)
            plan.launch()
            ctx.synchronize()
            for i, (v, (_, sub)) in enumerate(zip(views, hosts)):
                r = read_view(v, pool_size, sub)
                lit = v.scene.host.depth < 1.0  # sky pixels are left untouched by contract
                assert np.array_equal(r["out"][lit], eager[i]["out"][lit]), (rep, i)
                if v.adjoint:
                    assert close(r["dpx"], eager[i]["dpx"]) and close(r["dpos"], eager[i]["dpos"])
                    assert abs(r["loss"] - eager[i]["loss"]) <= 1e-6 * abs(eager[i]["loss"])
    finally:
        plan.destroy()


def metallic_f0_scene(sc):
    """A copy of sc with its G-buffer decoded to RGBA32F and the specular F0
    channels (normal.w, orm.w, emissive.w) set to the metallic-workflow F0
    the adjoint modes differentiate, clamp(0.04 (1 - metallic) + albedo
    metallic, 0, 0.999): the forward they compute.  (The forward reads F0 as
    stored: an RGBA8 orm.w holds a dielectric's 0.04 as 10/255.)"""
    def dec(a, fmt):
        if fmt == abi.HL_FMT_RGBA8_UNORM:
            return a.astype(np.float32) / np.float32(255.0)
        if fmt == abi.HL_FMT_RGBA16_FLOAT:
            return a.view(np.float16).astype(np.float32)
        return a.view(np.float32).copy()
    fa, _, fn, _ = sc.formats
    alb, orm = dec(sc.albedo, fa), dec(sc.orm, fa)
    nrm, emi = dec(sc.normal, fn), dec(sc.emissive, fn)
    m = orm[..., 2]
    f0 = [np.clip(np.float32(0.04) * (np.float32(1.0) - m) + alb[..., c] * m, np.float32(0.0), np.float32(0.999))
          for c in range(3)]
    nrm[..., 3], orm[..., 3], emi[..., 3] = f0
    f32 = abi.HL_FMT_RGBA32_FLOAT
    return dataclasses.replace(sc, albedo=np.ascontiguousarray(alb), orm=np.ascontiguousarray(orm),
                               normal=np.ascontiguousarray(nrm), emissive=np.ascontiguousarray(emi),
                               formats=(f32, 16, f32, 16))


def test_batch_views_against_the_oracle(ctx, hlib, mv):
    """Batched views against the oracle itself (not only the single-view
    path): probe faces of several light counts and both eyes, each against
    the oracle's own cluster cull and FP32 shading on the plan's grid --
    radiance within 1e-4 relative + 1e-6; the eyes' fused output against the
    forward on the metallic-workflow F0 it differentiates, and their L2 loss
    against that radiance's within 1e-4 relative."""
    import oracle_lib as O
    from test_gpu_parity import check_radiance
    pool, pool_size, views, hosts, _ = mv
    print("[batch] views against the oracle", flush=True)
    pick = sorted({0, 16, 31, 47} | {i for i, v in enumerate(views) if v.adjoint})
    plan = BatchPlan(ctx, pool, pool_size, views)
    try:
        for i in pick:
            views[i].out.fill_u32(0)
        plan.launch()
        ctx.synchronize()
        for i in pick:
            v, (sc, sub) = views[i], hosts[i]
            cfg = plan.config(i)
            params = synth.cluster_params(hlib, sc, grid=(cfg.grid_x, cfg.grid_y, cfg.grid_z), cap=0)
            off, cnt, opool, _ = O.cull_clusters(sc, params, 1 << 22)
            ref, _ = O.shade(metallic_f0_scene(sc) if v.adjoint else sc,
                             O.cluster_grid_host(params, off, cnt, opool))
            got = read_view(v, pool_size, sub)
            sky = sc.depth >= 1.0
            r = check_radiance(got["out"], ref, ~sky)
            print(f"  view {i} ({sc.width}x{sc.height}, {len(sub)} lights): max rel {r:.2e}")
            if v.adjoint:
                e = ref[..., :3].astype(np.float64) - 0.25
                oloss = float((e[~sky] ** 2).sum())
                assert abs(got["loss"] - oloss) <= 1e-4 * oloss, (got["loss"], oloss)
    finally:
        plan.destroy()


def test_autotuner_and_cache(ctx, mv, tmp_path):
    pool, pool_size, views, hosts, _ = mv
    print("[batch] autotune", flush=True)
    sel = views[::10] + [views[-1]]
    cache = str(tmp_path / "tune.txt")
    plan = BatchPlan(ctx, pool, pool_size, sel, autotune=True, max_trials=4, cache_path=cache)
    try:
        cfgs = [plan.config(i) for i in range(len(sel))]
        for c in cfgs:
            assert (c.wave_tile_w, c.wave_tile_h) in ((8, 8), (16, 4), (32, 2))
            assert 1 <= c.grid_z <= 64 and c.est_ms > 0
    finally:
        plan.destroy()
    text = open(cache).read()
    assert text.count("\n") >= 1
    plan2 = BatchPlan(ctx, pool, pool_size, sel, autotune=True, max_trials=4, cache_path=cache)
    try:
        for i, c in enumerate(cfgs):
            c2 = plan2.config(i)
            assert (c2.wave_tile_w, c2.grid_x, c2.grid_y, c2.grid_z) == (c.wave_tile_w, c.grid_x, c.grid_y, c.grid_z)
    finally:
        plan2.destroy()


def test_plan_rejects_bad_subset(ctx, mv):
    pool, pool_size, views, hosts, _ = mv
    v = views[0]
    bad = ctx.upload(np.array([5, 3, 9000], np.uint32))
    from helio_amd.batch import BatchView
    bv = BatchView(scene=v.scene, subset=bad, num_lights=3, out=v.out)
    with pytest.raises(abi.HelioError) as e:
        BatchPlan(ctx, pool, pool_size, [bv])
    assert e.value.status == abi.HL_EINVAL


def test_fixed_cluster_grid(ctx, hlib, mv):
    pool, pool_size, views, hosts, _ = mv
    print("[batch] fixed cluster grids", flush=True)
    idx = list(range(0, len(views), 7)) + [len(views) - 1]
    sel, fixed = [], []
    for k, i in enumerate(idx):
        sc = views[i].scene.host
        grid = ((sc.width + 31) // 32, (sc.height + 31) // 32, 16 if k % 2 else 40)
        cap = 24 if k % 3 == 0 else 0
        fixed.append((grid, cap))
        sel.append(dataclasses.replace(views[i], clusters=synth.cluster_params(hlib, sc, grid=grid, cap=cap)))
    plan = BatchPlan(ctx, pool, pool_size, sel, autotune=True, max_trials=2)
    try:
        for v in sel:
            v.out.fill_u32(0)
        plan.launch()
        ctx.synchronize()
        for k, (v, i) in enumerate(zip(sel, idx)):
            cfg = plan.config(k)
            assert (cfg.grid_x, cfg.grid_y, cfg.grid_z) == fixed[k][0]
            got = read_view(v, pool_size, hosts[i][1])
            ref = single_view(ctx, hlib, v, cfg, cap=fixed[k][1])
            sky = v.scene.host.depth >= 1.0
            if v.adjoint:
                assert close(got["out"][~sky], ref["out"][~sky], rel=1e-6), f"view {i}"
            else:
                assert np.array_equal(got["out"][~sky], ref["out"][~sky]), f"view {i}"
            if v.adjoint:
                assert close(got["dpx"], ref["dpx"]), f"view {i} pixel grads"
                assert close(got["dint"], ref["dint"]), f"view {i} intensity grads"
    finally:
        plan.destroy()


def test_fixed_cluster_grid_must_match_view(ctx, hlib, mv):
    pool, pool_size, views, hosts, _ = mv
    v = views[0]
    sc = v.scene.host
    p = synth.cluster_params(hlib, sc)
    p.width = sc.width - 1  # still covered by the grid, but not this view's size
    with pytest.raises(abi.HelioError, match="cluster grid is for"):
        BatchPlan(ctx, pool, pool_size, [dataclasses.replace(v, clusters=p)])


def test_full_size_c5(ctx, hlib):
    """Configuration 5 at full size in one autotuned plan: 48 probe faces + 2
    eyes at 2064x2208 (14.6 M pixels; the eyes' candidate boxes overflow
    mark-queue sub-queues on their finer grids).  Every view against the
    single-view path as in test_batch_equals_single_view (and the eyes' loss
    and gradients), one face of each size against the oracle in full, and
    three bands of rows of both eyes against the oracle (on the
    metallic-workflow F0 the fused adjoint differentiates)."""
    import oracle_lib as O
    from test_gpu_parity import check_radiance
    print("[batch] full size, one autotuned 50-view plan", flush=True)
    pool, pool_size, views, hosts, _ = make_multiview(ctx, small=False)
    for v in views:
        if v.adjoint:
            t = np.zeros((v.scene.host.height, v.scene.host.width, 4), np.float32)
            t[..., :3] = 0.25
            v.target.upload(t)
    plan = BatchPlan(ctx, pool, pool_size, views, autotune=True, max_trials=4)
    checked = set()
    try:
        plan.launch()
        ctx.synchronize()
        for i, v in enumerate(views):
            sc, sub = hosts[i]
            cfg = plan.config(i)
            got = read_view(v, pool_size, sub)
            ref = single_view(ctx, hlib, v, cfg)
            sky = sc.depth >= 1.0
            if v.adjoint and (cfg.wave_tile_w, cfg.wave_tile_h) != (8, 8):
                assert close(got["out"][~sky], ref["out"][~sky], rel=1e-6), f"view {i}"
            else:
                assert np.array_equal(got["out"][~sky], ref["out"][~sky]), f"view {i}"
            if v.adjoint:
                assert abs(got["loss"] - ref["loss"]) <= 1e-6 * abs(ref["loss"]), f"view {i} loss"
                assert close(got["dpx"], ref["dpx"]), f"view {i} pixel grads"
                assert close(got["dint"], ref["dint"]), f"view {i} intensity grads"
            key = "eye" if v.adjoint else (sc.width, sc.height)
            if key in checked and not v.adjoint:
                continue
            checked.add(key)
            params = synth.cluster_params(hlib, sc, grid=(cfg.grid_x, cfg.grid_y, cfg.grid_z), cap=0)
            off, cnt, opool, _ = O.cull_clusters(sc, params, 1 << 24)
            grid = O.cluster_grid_host(params, off, cnt, opool)
            if v.adjoint:
                osc = metallic_f0_scene(sc)
                for r0 in (0, sc.height // 2, sc.height - 32):
                    oref, _ = O.shade(osc, grid, rows=(r0, r0 + 32))
                    m = np.zeros_like(sky)
                    m[r0:r0 + 32] = True
                    r = check_radiance(got["out"], oref, m & ~sky)
                    print(f"  eye {i} rows {r0}..{r0 + 32}: max rel {r:.2e}", flush=True)
            else:
                oref, _ = O.shade(sc, grid)
                r = check_radiance(got["out"], oref, ~sky)
                print(f"  face {i} {(sc.width, sc.height)}: max rel {r:.2e}", flush=True)
    finally:
        plan.destroy()
    assert len(checked) == 4