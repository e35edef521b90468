"""GPU parity of the HIP path against the FP32 restatement (oracle).

Bars (stated per check):
  * light lists (tiles and clusters): bit-exact, every entry, every count;
  * radiance: |hip - oracle| <= 1e-4 * |oracle| + 1e-6 per channel, on the
    oracle evaluating the same lists;
  * Q (pixel-light pairs past the range test): equal.
"""
import ctypes as C
from dataclasses import replace

import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import (ClusterGrid, DeferredLightPass, LightCullPass, TileGrid,
                              light_cull_params, upload_scene)
import oracle_lib as O

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-4, 1e-6

_SCENES = {}


def scene(name):
    if name not in _SCENES:
        _SCENES[name] = synth.make_scene(name)
    return _SCENES[name]


def check_radiance(hip, ref, mask=None, rtol=RTOL, atol=ATOL):
    if mask is not None:
        hip, ref = hip[mask], ref[mask]
    err = np.abs(hip.astype(np.float64) - ref.astype(np.float64))
    tol = rtol * np.abs(ref.astype(np.float64)) + atol
    bad = err > tol
    # reported: the largest relative error where the absolute slack is under
    # 1 % of the bound (|ref| >= 1e-2)
    big = np.abs(ref) >= 1e-2
    rel = err[big] / np.abs(ref[big]).astype(np.float64)
    if bad.any():
        i = np.argwhere(bad)[:5]
        detail = "; ".join(f"{tuple(int(v) for v in k)} hip={hip[tuple(k)]:.8g} ref={ref[tuple(k)]:.8g}"
                           for k in i)
    else:
        detail = ""
    assert not bad.any(), (f"{int(bad.sum())} of {bad.size} values beyond 1e-4 rel; "
                           f"max rel {float(rel.max()) if rel.size else 0.0:.3e}: {detail}")
    return float(rel.max()) if rel.size else 0.0


def hip_tiles(ctx, sc):
    ds = upload_scene(ctx, sc)
    cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    cull.execute(ds)
    lists = cull.tiles.lists.download(np.uint32).reshape(-1, 64)
    counts = cull.tiles.counts.download(np.uint32)
    return ds, cull, lists, counts


def check_tile_lists(lists, counts, ol, oc):
    assert np.array_equal(counts, oc)
    for t in np.nonzero(oc)[0]:
        n = oc[t]
        assert np.array_equal(lists[t, :n], ol[t, :n]), f"tile {t}"


@pytest.mark.parametrize("name", ["c1_720p_64", "c2_1080p_1k", "c3_4k_16k", "c3_4k_16k_far", "around_camera"])
def test_tile_cull_bit_exact(ctx, name):
    """The screen's tile grid takes the binned path (candidates per 4 x 4
    super-tile, exact plane test per tile): the same lists as the reference's
    brute force, including lights behind and beside the camera."""
    if name == "around_camera":
        from test_tile_candidates import around_camera
        sc = around_camera(3, n=64, width=1280, height=720)
    else:
        sc = scene(name)
    print(f"[tile cull] {name}", flush=True)
    _, _, lists, counts = hip_tiles(ctx, sc)
    ol, oc = O.cull_tiles(sc)
    check_tile_lists(lists, counts, ol, oc)
    print(f"  tiles {len(oc)}, full tiles {float(np.mean(oc == 64)):.2f}")


def test_tile_cull_other_grid(ctx):
    """A tile grid other than the screen's (one column and row more) takes the
    brute-force kernel: the same lists as the oracle's for the same params."""
    import ctypes as C
    sc = scene("c2_1080p_1k")
    print("[tile cull] a grid wider than the screen", flush=True)
    p = light_cull_params(sc)
    p.num_tiles_x += 1
    p.num_tiles_y += 1
    tiles = p.num_tiles_x * p.num_tiles_y
    ds = upload_scene(ctx, sc)
    lists_d, counts_d = ctx.alloc(4 * 64 * tiles), ctx.alloc(4 * tiles)
    ctx.lib.call("hl_light_cull_tiles", ctx.h, ds.camera.ptr, C.byref(p), ds.lights.ptr, lists_d.ptr, counts_d.ptr)
    lists = lists_d.download(np.uint32).reshape(tiles, 64)
    counts = counts_d.download(np.uint32)
    ol = np.zeros(tiles * 64, np.uint32)
    oc = np.zeros(tiles, np.uint32)
    O.oracle().or_light_cull_tiles(C.byref(sc.camera), C.byref(p), sc.lights.ctypes.data, ol.ctypes.data,
                                   oc.ctypes.data)
    check_tile_lists(lists, counts, ol.reshape(tiles, 64), oc)


def hip_clusters(ctx, sc, params, pool_cap):
    ds = upload_scene(ctx, sc)
    cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=pool_cap)
    cull.prepare(ds)
    cull.execute(ds)
    off, cnt, pool = cull.cgrid.download()
    return ds, cull, off, cnt, pool, cull.overflow()


# Grids of at most 2^22 cells x lights take the brute-force launches, larger
# ones the light-parallel chain: the 1080p/1k scene runs on both.
@pytest.mark.parametrize("name,pool_cap,grid", [("c1_720p_64", 1 << 20, None), ("c2_1080p_1k", 1 << 21, None),
                                                ("c2_1080p_1k", 1 << 22, (48, 27, 24)),
                                                ("c4_512_adj", 1 << 20, None),
                                                ("c3_4k_16k", 1 << 25, None), ("c3_4k_16k_far", 1 << 25, None)])
def test_cluster_cull_bit_exact(ctx, hlib, name, pool_cap, grid):
    sc = scene(name)
    print(f"[cluster cull] {name} grid={grid}", flush=True)
    params = synth.cluster_params(hlib, sc) if grid is None else synth.cluster_params(hlib, sc, grid=grid)
    _, _, off, cnt, pool, flags = hip_clusters(ctx, sc, params, pool_cap)
    ooff, ocnt, opool, oflags = O.cull_clusters(sc, params, pool_cap)
    assert flags == oflags
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(off, ooff)
    total = int(ocnt.sum())
    # pool is dense in cluster order: compare the used prefix
    used = int(ooff[-1] + ocnt[-1]) if len(ocnt) else 0
    assert np.array_equal(pool[:used], opool[:used])
    print(f"  clusters={len(cnt)} entries={total} max={int(ocnt.max())} flags={flags}")
    if name == "c3_4k_16k_far":
        assert int(ocnt.max()) >= 2000  # the long-list regime


def test_cluster_cull_mark_queue_overflow(ctx, hlib):
    """128 large lights in front of the camera on a 64x36x24 grid (light-parallel
    path): the mark kernel queues each light's chunks past the first two in
    sub-queue (light % 64) of (4 L + 4096) / 64 entries, and here both lights of
    at least 32 sub-queues have more queued chunks than that, so the second
    to claim finds its sub-queue already full and must mark its chunks itself."""
    s = replace(synth.SPECS["c2_1080p_1k"], name="c2_overflow", num_lights=128,
                light_box_center=(0.0, 10.0, 40.0), light_box_size=(20.0, 6.0, 10.0),
                range_min=12.0, range_max=16.0)
    sc = synth.make_scene(s)
    params = synth.cluster_params(hlib, sc, grid=(64, 36, 24))
    print("[cluster cull] mark queue overflow 1920x1080 grid=(64,36,24) lights=128", flush=True)
    ooff, ocnt, opool, oflags = O.cull_clusters(sc, params, 1 << 22)
    per = np.bincount(opool[:int(ocnt.sum())], minlength=128)
    seg = (4 * 128 + 4096 + 63) // 64
    full = sum(bool(per[i] > 64 * (seg + 2) and per[i + 64] > 64 * (seg + 2)) for i in range(64))
    assert 64 * 36 * 24 * 128 > 1 << 22 and full >= 32
    _, _, off, cnt, pool, flags = hip_clusters(ctx, sc, params, 1 << 22)
    assert flags == oflags
    assert np.array_equal(cnt, ocnt) and np.array_equal(off, ooff)
    used = int(ooff[-1] + ocnt[-1])
    assert np.array_equal(pool[:used], opool[:used])
    print(f"  sub-queues over full={full} entries={used}")


@pytest.mark.parametrize("grid", [None, (48, 27, 24)])
def test_cluster_cap_and_pool_truncation(ctx, hlib, grid):
    sc = scene("c2_1080p_1k")
    print(f"[cluster truncation] grid={grid}", flush=True)
    params = (synth.cluster_params(hlib, sc, cap=32) if grid is None
              else synth.cluster_params(hlib, sc, grid=grid, cap=32))
    _, _, off, cnt, pool, flags = hip_clusters(ctx, sc, params, 5000)
    ooff, ocnt, opool, oflags = O.cull_clusters(sc, params, 5000)
    assert flags == oflags and (flags & abi.HL_OVERFLOW_CAP) and (flags & abi.HL_OVERFLOW_POOL)
    assert np.array_equal(cnt, ocnt) and np.array_equal(off, ooff)
    assert np.array_equal(pool, opool[:5000])


@pytest.mark.parametrize("name", ["c1_720p_64", "c2_1080p_1k"])
def test_shade_tiles_parity(ctx, name):
    sc = scene(name)
    print(f"[shade tiles] {name}", flush=True)
    ds, cull, lists, counts = hip_tiles(ctx, sc)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish())
    out = shade.output()
    pairs = shade.count_pairs(ds, cull.publish())
    lists = np.ascontiguousarray(lists.reshape(-1))
    grid = O.tile_grid_host(sc, lists, counts)
    ref, opairs = O.shade(sc, grid)
    sky = sc.depth >= 1.0
    assert np.all(out[sky] == 0.0)
    r = check_radiance(out, ref, ~sky)
    assert pairs == opairs
    print(f"  max rel err {r:.2e}, Q={pairs}")


@pytest.mark.parametrize("name,rows", [("c1_720p_64", None), ("c2_1080p_1k", None),
                                       ("c3_4k_16k", None), ("c3_4k_16k_far", None)])
def test_shade_clusters_parity(ctx, hlib, name, rows):
    """Every pixel of the frame, 4K included; c3_4k_16k_far walks lists of up
    to 2000+ lights."""
    sc = scene(name)
    print(f"[shade clusters] {name}", flush=True)
    params = synth.cluster_params(hlib, sc)
    ds, cull, off, cnt, pool, flags = hip_clusters(ctx, sc, params, 1 << 25)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish())
    out = shade.output()
    grid = O.cluster_grid_host(params, off, cnt, pool)
    ref, opairs = O.shade(sc, grid, rows=rows)
    sl = slice(None) if rows is None else slice(rows[0], rows[1])
    sky = sc.depth[sl] >= 1.0
    r = check_radiance(out[sl], ref[sl], ~sky)
    if rows is None:
        assert shade.count_pairs(ds, cull.publish()) == opairs
    print(f"  max rel err {r:.2e}")


@pytest.mark.parametrize("name,over", [("c4_512_adj", {}),
                                       ("c3_4k_16k", dict(width=960, height=540))])
def test_smooth_surfaces_near_the_specular_peak(ctx, hlib, name, over):
    """Roughness 0.03-0.25 on every pixel: many terms lie near the GGX peak,
    where the denominator NdH^2 (a^2 - 1) + 1 is a few 1e-3 and a last-bit
    difference of L or H would move D by percents; every pixel within 1e-4 of
    the oracle (fp32 and fp16 G-buffers)."""
    sc = synth.make_scene(name, **over)
    dt = {abi.HL_FMT_RGBA32_FLOAT: np.float32, abi.HL_FMT_RGBA16_FLOAT: np.float16}[sc.formats[0]]
    orm = sc.orm.view(dt).reshape(sc.height, sc.width, 4)
    orm[..., 1] = np.random.default_rng(7).uniform(0.03, 0.25, (sc.height, sc.width)).astype(dt)
    print(f"[smooth] {name} {sc.width}x{sc.height}", flush=True)
    params = synth.cluster_params(hlib, sc)
    ds, cull, off, cnt, pool, flags = hip_clusters(ctx, sc, params, 1 << 25)
    assert flags == 0
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish())
    ref, pairs = O.shade(sc, O.cluster_grid_host(params, off, cnt, pool))
    r = check_radiance(shade.output(), ref, sc.depth < 1.0)
    print(f"  max rel err {r:.2e}, Q={pairs}")


def test_zero_lights(ctx):
    sc = synth.make_scene("c1_720p_64", num_lights=0)
    ds = upload_scene(ctx, sc)
    cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    cull.tiles.counts.fill_u32(7)
    cull.execute(ds)
    assert np.all(cull.tiles.counts.download(np.uint32) == 0)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.execute(ds, cull.publish(), flags=abi.HL_SHADE_CLEAR_SKY)
    out = shade.output()
    lists = np.zeros(cull.tiles.ntx * cull.tiles.nty * 64, np.uint32)
    counts = np.zeros(cull.tiles.ntx * cull.tiles.nty, np.uint32)
    ref, pairs = O.shade(sc, O.tile_grid_host(sc, lists, counts), flags=abi.HL_SHADE_CLEAR_SKY)
    assert pairs == 0
    check_radiance(out, ref)


def test_shade_lights_from_cull(ctx):
    """HL_SHADE_LIGHTS_FROM_CULL: shading with the records the cluster cull
    built equals shading that builds them (bit for bit); it is refused after a
    cull over other lights."""
    sc = scene("c1_720p_64")
    ds = upload_scene(ctx, sc)
    params = synth.cluster_params(ctx.lib, sc)
    cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 20)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    shade.prepare(ds)
    cull.execute(ds)
    grid = cull.publish()
    shade.execute(ds, grid)
    ref = shade.output().copy()
    cull.execute(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, grid, flags=abi.HL_SHADE_LIGHTS_FROM_CULL)
    got = shade.output()
    lit = sc.depth < 1.0
    assert np.array_equal(got[lit], ref[lit])
    # a shading call that builds records for other lights, then reuse: refused
    few = synth.make_scene("c1_720p_64", num_lights=3)
    ds_few = upload_scene(ctx, few)
    shade.execute(ds_few, grid)
    with pytest.raises(abi.HelioError, match="LIGHTS_FROM_CULL"):
        shade.execute(ds, grid, flags=abi.HL_SHADE_LIGHTS_FROM_CULL)


@pytest.mark.parametrize("mode", [1, 2, 4, 5, 20, 21])
def test_debug_views_parity(ctx, mode):
    """Debug views of fs_main (deferred_lighting.wgsl:1032-1038): albedo for
    1/2/4/20/21 (bit-exact decode), the normal mapped to [0,1] for 5 (1e-4)."""
    sc = scene("c1_720p_64")
    print(f"[debug view] {mode}", flush=True)
    ds, cull, lists, counts = hip_tiles(ctx, sc)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.debug_mode = mode
    shade.prepare(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish())
    out = shade.output()
    grid = O.tile_grid_host(sc, np.ascontiguousarray(lists.reshape(-1)), counts)
    ref, _ = O.shade(sc, grid, globals_=shade.globals)
    lit = sc.depth < 1.0
    if mode == 5:
        check_radiance(out, ref, lit)
    else:
        assert np.array_equal(out[lit], ref[lit])


def test_unsorted_slice_tables_are_rejected(ctx, hlib):
    """The cull searches slice_near / slice_far as sorted tables under the
    padding of the farthest slice in reach: tables that are not
    non-decreasing, or a negative padding, are refused."""
    sc = scene("c1_720p_64")
    ds = upload_scene(ctx, sc)
    for field, k, value in (("slice_near", 3, None), ("slice_far", 5, None), ("slice_r_pad", 2, -1.0)):
        p = synth.cluster_params(hlib, sc)
        arr = getattr(p, field)
        arr[k] = arr[k - 1] * 0.5 if value is None else value
        with pytest.raises(abi.HelioError, match="slice"):
            cull = LightCullPass(ctx, sc.width, sc.height, clusters=p, pool_cap=1 << 20)
            cull.prepare(ds)
            cull.execute(ds)


def test_unsupported_debug_view_is_rejected(ctx):
    sc = scene("c1_720p_64")
    ds, cull, _, _ = hip_tiles(ctx, sc)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.debug_mode = 3  # needs the shadow atlas
    shade.prepare(ds)
    with pytest.raises(abi.HelioError) as e:
        shade.execute(ds, cull.publish())
    assert e.value.status == abi.HL_EUNSUPPORTED


@pytest.mark.parametrize("name,shadowed", [("c3_4k_16k", False), ("c2_1080p_1k", True)])
def test_rgba16f_target(ctx, hlib, name, shadowed):
    """HL_SHADE_OUT_RGBA16F (the pass's pre_aa format): every pixel of the
    RGBA16F target is the same call's RGBA32F result rounded to the nearest
    binary16 (ties to even), bit for bit, cleared sky included."""
    from helio_amd.device import DeviceShadows
    print(f"[rgba16f] {name} shadowed={shadowed}", flush=True)
    sc = synth.make_scene(name)
    sh = synth.add_shadow_casters(sc, n_point=4, n_spot=4, size=256) if shadowed else None
    params = synth.cluster_params(hlib, sc)
    ds, cull, off, cnt, pool, flags = hip_clusters(ctx, sc, params, 1 << 25)
    dsh = DeviceShadows(ctx, sh) if shadowed else None
    outs = []
    for fmt in (abi.HL_FMT_RGBA32_FLOAT, abi.HL_FMT_RGBA16_FLOAT):
        shade = DeferredLightPass(ctx, sc.width, sc.height, out_format=fmt)
        shade.prepare(ds)
        shade.out.fill_u32(0x7FC00000)  # a NaN pattern where nothing is written
        shade.execute(ds, cull.publish(), flags=abi.HL_SHADE_CLEAR_SKY, shadows=dsh)
        outs.append(shade.out.download(np.float32 if fmt == abi.HL_FMT_RGBA32_FLOAT else np.float16,
                                       (sc.height, sc.width, 4)))
    with np.errstate(over="ignore"):
        want = outs[0].astype(np.float16)
    assert np.array_equal(outs[1].view(np.uint16), want.view(np.uint16))
    assert np.all(outs[1][sc.depth >= 1.0] == 0)


def test_walk_stats(ctx):
    """hl_shade_walk_stats: with directional lights only (every tile lists
    them and they pass every bounds test) each 8x8 wavefront with a shaded
    pixel visits each light once, and Q is lights x shaded pixels; with point
    lights, Q equals the oracle's and 64 visits cover Q."""
    print("[walk stats]", flush=True)
    sc = synth.make_scene("c1_720p_64", num_lights=5)
    d = np.array([0.35, -1.0, 0.25]) / np.linalg.norm([0.35, -1.0, 0.25])
    for i in range(5):
        sc.lights[i]["light_type"] = abi.HL_LIGHT_DIRECTIONAL
        sc.lights[i]["direction_outer"] = (*d, 0.0)
    ds, cull, _, _ = hip_tiles(ctx, sc)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    q, visits = shade.walk_stats(ds, cull.publish())
    valid = sc.depth < 1.0
    h8, w8 = (sc.height + 7) // 8, (sc.width + 7) // 8
    pad = np.zeros((h8 * 8, w8 * 8), bool)
    pad[:sc.height, :sc.width] = valid
    waves = int(pad.reshape(h8, 8, w8, 8).any(axis=(1, 3)).sum())
    assert q == 5 * int(valid.sum())
    assert visits == 5 * waves

    sc = scene("c1_720p_64")
    ds, cull, lists, counts = hip_tiles(ctx, sc)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    q, visits = shade.walk_stats(ds, cull.publish())
    assert q == shade.count_pairs(ds, cull.publish())
    _, opairs = O.shade(sc, O.tile_grid_host(sc, np.ascontiguousarray(lists.reshape(-1)), counts))
    assert q == opairs
    assert 64 * visits >= q > 0
    print(f"  Q {q}, visits {visits}, lane utilisation {q / (64 * visits):.3f}")
