"""The interop bridge (hl_external_import / hl_external_map /
hl_external_release): device memory exported by another allocator as a
file descriptor -- a HIP virtual-memory allocation here, in place of an
exportable Vulkan VkDeviceMemory -- written by the light cull and the
deferred forward through the imported mapping and read back through the
exporter's own mapping.

Bars: the tile lists written through the import equal the oracle's bit for
bit; the radiance written through the import equals the same call's output
into library-allocated memory bit for bit (RGBA32F and RGBA16F targets);
bytes outside a mapped range are untouched."""
import ctypes as C
import os
import types

import numpy as np
import pytest

from helio_amd import abi
from helio_amd.device import Context, DeferredLightPass, ExternalMemory, LightCullPass, upload_scene
import oracle_lib as O
from hip_vmm import ExportedAllocation
from test_gpu_parity import check_tile_lists, scene

pytestmark = pytest.mark.gpu

ALIGN = abi.HL_EXTERNAL_MAP_ALIGN


def read(ctx, ptr, nbytes, dtype):
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    ctx.lib.call("hl_memcpy_dtoh", ctx.h, out.ctypes.data, ptr, out.nbytes)
    return out


def up(n):
    return (n + ALIGN - 1) // ALIGN * ALIGN


def test_round_trip_and_offsets(ctx):
    print("[interop] round trip", flush=True)
    with ExportedAllocation(8 << 20) as ea:
        ctx.lib.call("hl_memset_d32", ctx.h, ea.ptr, 0x12345678, ea.size // 4)
        with ExternalMemory(ctx, ea.fd, ea.size) as ext:
            off = 3 * ALIGN
            p = ext.map(off, 4096)
            ctx.lib.call("hl_memset_d32", ctx.h, p, 0xABCDEF01, 1024)
            whole = read(ctx, ea.ptr, ea.size, np.uint32)
            w0 = off // 4
            assert np.all(whole[w0:w0 + 1024] == 0xABCDEF01)
            assert np.all(whole[:w0] == 0x12345678) and np.all(whole[w0 + 1024:] == 0x12345678)
            # the exporter's writes are seen through the import
            ctx.lib.call("hl_memset_d32", ctx.h, ea.ptr + ea.size - 4096, 7, 1024)
            q = ext.map(ea.size - 4096, 4096)
            assert np.all(read(ctx, q, 4096, np.uint32) == 7)
        os.fstat(ea.fd)  # the caller's descriptor stays open after the release


@pytest.mark.parametrize("name,fmt", [("c1_720p_64", abi.HL_FMT_RGBA32_FLOAT),
                                      ("c3_4k_16k", abi.HL_FMT_RGBA16_FLOAT)])
def test_cull_and_forward_through_import(ctx, name, fmt):
    """LightCull's tile_light_lists / tile_light_counts and DeferredLight's
    pre_aa (crates/libhelio/src/frame.rs:228-234) in one imported allocation."""
    print(f"[interop] {name} fmt={fmt}", flush=True)
    sc = scene(name)
    ds = upload_scene(ctx, sc)
    cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    ntiles = cull.tiles.ntx * cull.tiles.nty
    lists_b, counts_b = ntiles * abi.HL_MAX_LIGHTS_PER_TILE * 4, ntiles * 4
    shade = DeferredLightPass(ctx, sc.width, sc.height, out_format=fmt)
    shade.prepare(ds)
    out_b = shade.pitch * sc.height
    o_counts = up(lists_b)
    o_out = up(o_counts + counts_b)
    flags = abi.HL_SHADE_CLEAR_SKY
    # the same calls into library memory
    cull.execute(ds)
    shade.execute(ds, cull.publish(), flags=flags)
    want_out = shade.out.download(np.uint8)[:out_b]
    with ExportedAllocation(o_out + out_b) as ea, ExternalMemory(ctx, ea.fd, ea.size) as ext:
        ctx.lib.call("hl_memset_d32", ctx.h, ea.ptr, 0xFFFFFFFF, ea.size // 4)
        lists_d, counts_d, out_d = ext.map(0, lists_b), ext.map(o_counts, counts_b), ext.map(o_out, out_b)
        ctx.lib.call("hl_light_cull_tiles", ctx.h, ds.camera.ptr, C.byref(cull.params), ds.lights.ptr, lists_d,
                     counts_d)
        grid = cull.tiles.grid()
        grid.lists_d, grid.counts_d = lists_d, counts_d
        lib_out = shade.out
        shade.out = types.SimpleNamespace(ptr=out_d)
        try:
            shade.execute(ds, grid, flags=flags)
        finally:
            shade.out = lib_out
        raw = read(ctx, ea.ptr, ea.size, np.uint8)
    lists = raw[:lists_b].view(np.uint32).reshape(-1, abi.HL_MAX_LIGHTS_PER_TILE)
    counts = raw[o_counts:o_counts + counts_b].view(np.uint32)
    ol, oc = O.cull_tiles(sc)
    check_tile_lists(lists, counts, ol, oc)
    assert np.array_equal(raw[o_out:o_out + out_b], want_out)
    # nothing outside the mapped ranges was written
    assert np.all(raw[lists_b:o_counts] == 0xFF) and np.all(raw[o_counts + counts_b:o_out] == 0xFF)
    assert np.all(raw[o_out + out_b:] == 0xFF)


def test_argument_errors(ctx):
    print("[interop] argument errors", flush=True)
    L = ctx.lib.raw
    h = C.c_void_p()

    def imp(fd, size, kind=abi.HL_EXTERNAL_FD):
        return L["hl_external_import"](ctx.h, fd, size, kind, C.byref(h))

    assert imp(-1, 4096) == abi.HL_EINVAL
    r, w = os.pipe()
    try:
        assert imp(r, 4096) == abi.HL_EINVAL  # no size: not a dma-buf
        assert "no size" in ctx.last_error()
    finally:
        os.close(r)
        os.close(w)
    with ExportedAllocation(1 << 20) as ea:
        assert imp(ea.fd, 0) == abi.HL_EINVAL
        assert imp(ea.fd, ea.size + 4096) == abi.HL_EINVAL
        assert "from a descriptor of" in ctx.last_error()
        assert imp(ea.fd, ea.size, kind=2) == abi.HL_EUNSUPPORTED
        assert L["hl_external_import"](ctx.h, ea.fd, ea.size, abi.HL_EXTERNAL_FD, None) == abi.HL_EINVAL
        with ExternalMemory(ctx, ea.fd, ea.size) as ext:
            p = C.c_void_p()
            m = L["hl_external_map"]
            assert m(ext.h, 0, ea.size + 1, C.byref(p)) == abi.HL_EINVAL
            assert m(ext.h, ea.size, ALIGN, C.byref(p)) == abi.HL_EINVAL
            assert m(ext.h, 0, 0, C.byref(p)) == abi.HL_EINVAL
            assert m(ext.h, ALIGN // 2, ALIGN, C.byref(p)) == abi.HL_EINVAL  # offset alignment
            assert m(ext.h, (1 << 64) - ALIGN, 2 * ALIGN, C.byref(p)) == abi.HL_EINVAL  # wraps around
            assert m(ext.h, 0, ALIGN, None) == abi.HL_EINVAL
            assert p.value is None
            assert m(ext.h, ea.size - ALIGN, ALIGN, C.byref(p)) == abi.HL_OK and p.value
        assert L["hl_external_release"](None) == abi.HL_EINVAL
        assert L["hl_external_map"](None, 0, ALIGN, C.byref(C.c_void_p())) == abi.HL_EINVAL
        os.fstat(ea.fd)


def test_context_destroy_releases_imports():
    print("[interop] context destroy", flush=True)
    with ExportedAllocation(1 << 20) as ea:
        c = Context(0)
        ext = ExternalMemory(c, ea.fd, ea.size)
        p = ext.map(0, 4096)
        c.lib.call("hl_memset_d32", c.h, p, 5, 1024)
        c.close()  # releases the import it still holds
        ext.h = None
        os.fstat(ea.fd)
        with Context(0) as c2:
            assert np.all(read(c2, ea.ptr, 4096, np.uint32) == 5)
