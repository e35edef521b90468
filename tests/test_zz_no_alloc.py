"""hl_deferred_forward_noalloc (scratch in a fixed device array of the
library, no allocation) against hl_deferred_forward on the same inputs.

Bar: bit-identical output (the same kernels over the same lists), on two
calls in a row -- the second on a smaller image with fewer lights, so that
the array still holds the first call's records and counters -- plus the
limits: more than HL_NOALLOC_MAX_LIGHTS lights, or a flag other than
HL_SHADE_CLEAR_SKY, is HL_EINVAL.
"""
import ctypes as C

import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import DeferredLightPass, LightCullPass, upload_scene

pytestmark = pytest.mark.gpu


def shade_both(ctx, sc, clusters: bool):
    ds = upload_scene(ctx, sc)
    if clusters:
        cull = LightCullPass(ctx, sc.width, sc.height, clusters=synth.cluster_params(ctx.lib, sc), pool_cap=1 << 22)
    else:
        cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    cull.execute(ds)
    grid = cull.publish()
    ref = DeferredLightPass(ctx, sc.width, sc.height)
    ref.prepare(ds)
    out = ctx.alloc(sc.width * sc.height * 16)
    res = []
    for flags in (abi.HL_SHADE_CLEAR_SKY, 0):
        ref.out.fill_u32(0x7fc00000)
        out.fill_u32(0x7fc00000)
        ref.execute(ds, grid, flags=flags)
        ctx.lib.call("hl_deferred_forward_noalloc", ctx.h, ds.camera.ptr, C.byref(ref.globals), C.byref(ds.gbuffer),
                     ds.lights.ptr, ds.num_lights, C.byref(grid), out.ptr, sc.width * 16, flags, ctx=ctx.h)
        ctx.synchronize()
        a = ref.output().view(np.uint32)
        b = out.download(np.float32, (sc.height, sc.width, 4)).view(np.uint32)
        res.append((a, b))
    return ds, grid, ref, res


def test_noalloc_matches_forward_on_two_calls(ctx):
    print("[no-alloc] 640x360 / 1024 lights (clusters), then 320x180 / 64 lights (tiles)", flush=True)
    big = synth.make_scene("c2_1080p_1k", width=640, height=360)
    small = synth.make_scene("c1_720p_64", width=320, height=180)
    for sc, clusters in ((big, True), (small, False)):
        _, _, _, res = shade_both(ctx, sc, clusters)
        for a, b in res:
            assert np.array_equal(a, b), f"{sc.width}x{sc.height}: {int((a != b).sum())} words differ"


def test_noalloc_limits(ctx):
    sc = synth.make_scene("c1_720p_64", width=64, height=48)
    ds, grid, ref, _ = shade_both(ctx, sc, False)
    out = ctx.alloc(sc.width * sc.height * 16)
    with pytest.raises(abi.HelioError, match="HL_NOALLOC_MAX_LIGHTS") as e:
        ctx.lib.call("hl_deferred_forward_noalloc", ctx.h, ds.camera.ptr, C.byref(ref.globals), C.byref(ds.gbuffer),
                     ds.lights.ptr, abi.HL_NOALLOC_MAX_LIGHTS + 1, C.byref(grid), out.ptr, sc.width * 16, 0,
                     ctx=ctx.h)
    assert e.value.status == abi.HL_EINVAL
    with pytest.raises(abi.HelioError, match="HL_SHADE_CLEAR_SKY only") as e:
        ctx.lib.call("hl_deferred_forward_noalloc", ctx.h, ds.camera.ptr, C.byref(ref.globals), C.byref(ds.gbuffer),
                     ds.lights.ptr, ds.num_lights, C.byref(grid), out.ptr, sc.width * 16,
                     abi.HL_SHADE_LIGHTS_FROM_CULL, ctx=ctx.h)
    assert e.value.status == abi.HL_EINVAL
