"""Shadowed deferred shading and radiance-cascade GI on the GPU
(hl_deferred_forward_shadowed, hl_deferred_forward_ex) against
the FP32 restatement with shadow_factor (oracle/shadow.c) on the same lists:
radiance within 1e-4 relative (+1e-6 absolute).  The shadow factors are
reproducible bit for bit (own sin/cos, FP32 sampler weights), so a tap that
lands on another texel or a blocker counted differently would show as a
large error, not a small one."""
import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import DeferredLightPass, DeviceShadows, LightCullPass, upload_scene
import oracle_lib as O
from test_gpu_parity import check_radiance

pytestmark = pytest.mark.gpu


def run(ctx, hlib, name, mode, quality, size, flags=0, **over):
    sc = synth.make_scene(name, **over)
    sh = synth.add_shadow_casters(sc, n_point=8, n_spot=8, size=size, quality=quality)
    ds = upload_scene(ctx, sc)
    dsh = DeviceShadows(ctx, sh)
    if mode == "tiles":
        cull = LightCullPass(ctx, sc.width, sc.height)
    else:
        params = synth.cluster_params(hlib, sc)
        cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 22)
    cull.prepare(ds)
    cull.execute(ds)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish(), flags=flags, shadows=dsh)
    out = shade.output()
    if mode == "tiles":
        lists = cull.tiles.lists.download(np.uint32)
        counts = cull.tiles.counts.download(np.uint32)
        grid = O.tile_grid_host(sc, lists, counts)
        grid._arrays = (lists, counts)
    else:
        off, cnt, pool = cull.cgrid.download()
        grid = O.cluster_grid_host(params, off, cnt, pool)
    return sc, sh, out, grid, (ds, cull, shade, dsh)


# The 4K/16k-light scene at a quarter of its pixels and lights (fp16 G-buffer
# set) and the 512^2 scene (fp32 set) build the other two G-buffer variants.
C3_REDUCED = dict(width=960, height=540, num_lights=4096)


@pytest.mark.parametrize("name,mode,quality,size,over", [("c1_720p_64", "tiles", "medium", 256, {}),
                                                         ("c1_720p_64", "clusters", "high", 256, {}),
                                                         ("c2_1080p_1k", "tiles", "ultra", 512, {}),
                                                         ("c2_1080p_1k", "clusters", "low", 1024, {}),
                                                         ("c3_4k_16k", "clusters", "medium", 512, C3_REDUCED),
                                                         ("c4_512_adj", "tiles", "high", 256, {})])
def test_shadowed_shading_matches_oracle(ctx, hlib, name, mode, quality, size, over):
    print(f"[shadowed] {name} {mode} {quality} atlas {size} {over}", flush=True)
    sc, sh, out, grid, _ = run(ctx, hlib, name, mode, quality, size, **over)
    ref, _ = O.shade(sc, grid, shadows=sh)
    unshadowed, _ = O.shade(sc, grid)
    sky = sc.depth >= 1.0
    r = check_radiance(out, ref, ~sky)
    dark = np.mean(np.any(ref[~sky][:, :3] < unshadowed[~sky][:, :3] - 1e-6, axis=1))
    print(f"  max rel err {r:.2e}, shadowed pixels {dark:.2f}")
    assert dark > 0.05


def test_shadowed_with_cull_records(ctx, hlib):
    """HL_SHADE_LIGHTS_FROM_CULL: the cull's records carry the same shadow
    words, so the image is the same bits."""
    print("[shadowed] lights from cull", flush=True)
    sc, sh, out, grid, (ds, cull, shade, dsh) = run(ctx, hlib, "c1_720p_64", "clusters", "medium", 256)
    cull.execute(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish(), flags=abi.HL_SHADE_LIGHTS_FROM_CULL, shadows=dsh)
    assert np.array_equal(shade.output(), out)


def test_shadow_layer_past_the_atlas_is_lit(ctx, hlib):
    print("[shadowed] shadow layers past the atlas", flush=True)
    sc = synth.make_scene("c1_720p_64", width=320, height=180)
    sh = synth.add_shadow_casters(sc, n_point=4, n_spot=4, size=64)
    sc.lights["shadow_index"][1:3] = [sh.layers + 2, 0x7FFFFFFF]
    ds = upload_scene(ctx, sc)
    dsh = DeviceShadows(ctx, sh)
    cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    cull.execute(ds)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.execute(ds, cull.publish(), shadows=dsh)
    lists = cull.tiles.lists.download(np.uint32)
    counts = cull.tiles.counts.download(np.uint32)
    grid = O.tile_grid_host(sc, lists, counts)
    ref, _ = O.shade(sc, grid, shadows=sh)
    check_radiance(shade.output(), ref, sc.depth < 1.0)


@pytest.mark.parametrize("name,mode,fmt,shadows,over", [
    ("c1_720p_64", "tiles", "f16", False, {}),
    ("c2_1080p_1k", "clusters", "f32", True, {}),
    ("c3_4k_16k", "clusters", "f16", True, C3_REDUCED),
    ("c4_512_adj", "tiles", "f32", False, {})])
def test_rc_gi_matches_oracle(ctx, hlib, name, mode, fmt, shadows, over):
    """hl_deferred_forward_ex with the radiance-cascade indirect diffuse (and
    shadows): the RC irradiance is exact, the combine within 1e-4."""
    from helio_amd.device import DeviceTexture
    print(f"[rc gi] {name} {mode} {fmt} shadows={shadows} {over}", flush=True)
    sc = synth.make_scene(name, **over)
    sh = synth.add_shadow_casters(sc, n_point=4, n_spot=4, size=256) if shadows else None
    box = ((-60.0, -1.0, -60.0), (60.0, 30.0, 60.0)) if name == "c3_4k_16k" else ((-25.0, -1.0, -25.0),
                                                                                  (25.0, 14.0, 25.0))
    atlas = synth.make_rc_probes(sc, fmt=abi.HL_FMT_RGBA16_FLOAT if fmt == "f16" else abi.HL_FMT_RGBA32_FLOAT,
                                 box=box)
    ds = upload_scene(ctx, sc)
    if mode == "tiles":
        cull = LightCullPass(ctx, sc.width, sc.height)
    else:
        params = synth.cluster_params(hlib, sc)
        cull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 22)
    cull.prepare(ds)
    cull.execute(ds)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.execute(ds, cull.publish(), shadows=DeviceShadows(ctx, sh) if shadows else None,
                  rc=DeviceTexture(ctx, atlas))
    out = shade.output()
    if mode == "tiles":
        lists = cull.tiles.lists.download(np.uint32)
        counts = cull.tiles.counts.download(np.uint32)
        grid = O.tile_grid_host(sc, lists, counts)
        grid._arrays = (lists, counts)
    else:
        off, cnt, pool = cull.cgrid.download()
        grid = O.cluster_grid_host(params, off, cnt, pool)
    ref, _ = O.shade(sc, grid, shadows=sh, rc=atlas)
    plain, _ = O.shade(sc, grid, shadows=sh)
    sky = sc.depth >= 1.0
    r = check_radiance(out, ref, ~sky)
    changed = np.mean(np.any(ref[~sky][:, :3] != plain[~sky][:, :3], axis=1))
    print(f"  max rel err {r:.2e}, pixels with RC ambient {changed:.2f}")
    assert changed > 0.1


def test_rc_ambient_matches_float64(ctx):
    """The kernel's radiance-cascade ambient (k_rc_irradiance, then the
    extended shading launch) against the float64 restatement of
    sample_rc_irradiance and the combine (tests/rc64.py), on a frame without
    lights and random probes: within 2e-5 relative plus 2^-17 of the atlas's
    largest value, as the oracle (test_rc_gi.py)."""
    import rc64
    from helio_amd.device import DeviceTexture
    print("[rc gi] kernel ambient against float64", flush=True)
    sc = synth.make_scene("c4_512_adj", width=128, height=128, num_lights=0)
    lo, hi = (-40.0, -5.0, -40.0), (40.0, 25.0, 40.0)
    atlas = synth.make_rc_probes(sc, fmt=abi.HL_FMT_RGBA32_FLOAT, box=(lo, hi))
    ds = upload_scene(ctx, sc)
    cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    cull.execute(ds)
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.execute(ds, cull.publish(), rc=DeviceTexture(ctx, atlas))
    out = shade.output()
    ys, xs = np.nonzero(sc.depth < 1.0)
    want, cw = rc64.ambient64(sc, atlas, lo, hi, xs, ys)
    keep = ~np.any(np.abs(cw - 0.001) < 1e-6, axis=1)
    amax = np.abs(atlas[..., :3]).max()
    err = np.abs(out[ys, xs, :3] - want)[keep]
    assert keep.sum() > 5000
    bound = 2e-5 * np.abs(want[keep]) + 2.0 ** -17 * amax
    print(f"  max err / bound {float((err / bound).max()):.3f}")
    assert np.all(err <= bound)
